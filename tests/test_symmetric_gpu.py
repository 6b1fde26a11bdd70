"""The reference's single-solve path for rotation-symmetric targets
(DDA/DDAMain.cpp:380-410, 452-470) through dda_amd.solve_particle: a lattice
sphere lit along z is solved once; its polarisation-2 cross sections and its
Jones matrices (the rotated polarisation-1 solution, FarFieldCalc.cpp:167-168)
equal those of solving both polarisations, to the solver's tolerance; other
targets and directions solve both."""
import math

import numpy as np
import pytest

import dda_amd as D
from dda_amd import output as O
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_symmetric_single_solve_matches_both_solves(ctx, mode):
    g, params, cc, cs, prop, pol = cases.build("sphere_r7")
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid)
    s = D.GpuQmrCs(op, mode=mode, max_iter=5000)
    eps = 1e-10 if mode == "fp64" else 1e-8
    one = D.solve_particle(s, params, (0, 0, 1), eps)
    assert one["symmetric"] and one["nrhs"] == 1
    both = D.solve_particle(s, params, (0, 0, 1), eps, symmetric=False)
    assert not both["symmetric"] and both["nrhs"] == 2
    tol = 1e-8 if mode == "fp64" else 1e-6
    np.testing.assert_allclose(one["cext"], both["cext"], rtol=tol)
    np.testing.assert_allclose(one["cabs"], both["cabs"], rtol=tol)
    args = (one["prop"], one["pol1"], one["pol2"], 91, 0.0, math.radians(2), 61, 0.0, math.radians(6))
    J1 = D.jones(op, one["pol_dev"], 1, *args)
    J2 = D.jones(op, both["pol_dev"], 0, *args)
    d = O.field_diff(J1, J2)
    print(mode, "Jones FieldDiff symmetric vs both solves %.2e" % d, flush=True)
    assert d < tol
    for r in (one, both):
        r["pol_dev"].close()
    # not symmetric: a box along z, the sphere off-axis
    gb, pb, _, _, _, _ = cases.build("box_10_8_6")
    opb = D.GpuMatVec(ctx, gb.pos, gb.mat, gb.box, gb.grid_unit, pb.wave_num, grid=pb.grid)
    sb = D.GpuQmrCs(opb, mode=mode, max_iter=5000)
    rb = D.solve_particle(sb, pb, (0, 0, 1), eps)
    assert not rb["symmetric"] and rb["nrhs"] == 2 and all(i["status"] == D.OK for i in rb["infos"])
    ro = D.solve_particle(s, params, (0.6, 0, 0.8), eps)
    assert not ro["symmetric"] and ro["nrhs"] == 2
    for r in (rb, ro):
        r["pol_dev"].close()
    sb.close()
    opb.close()
    s.close()
    op.close()
