"""FP32-inner / FP64-outer refinement: same answer as the FP64 oracle to the
requested residual, Cext within 1e-5 of the FP64 solve."""
import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu


def _op(ctx, g, params, cs):
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=64, grid=params.grid,
                     nmat=len(g.materials))
    op.set_couple_constants(cs)
    return op


@pytest.mark.parametrize("kind", ["qmr", "bicgstab", "cgnr", "bicg"])
@pytest.mark.parametrize("name", ["sphere_r7", "two_materials"])
def test_mixed_matches_fp64_oracle(ctx, kind, name):
    print("mixed", kind, name, flush=True)
    g, params, cc, cs, prop, pol = cases.build(name)
    einc = P.einc_plane(params, prop, pol)
    orc = cases.oracle_problem(g, params, cs)
    st, p_ref, it_ref, _ = orc.solve(kind, einc, 1e-11, max_iter=5000)
    assert st == 0
    s = D.SOLVER_CLASSES[kind](_op(ctx, g, params, cs), max_iter=5000, mixed=True)
    eps = 1e-6
    p = s.get_pol_vec(einc, eps)
    info = s.info
    assert info["true_rel_residual"] < eps
    assert info["refinements"] >= 1
    # the residual bound eps limits the solution error to about eps * cond(A)
    assert cases.rel_err(p, p_ref) <= 1e-4
    cext, cext_ref = P.ext_cross(params, p, einc), orc.ext_cross_sum(p_ref, einc)
    assert abs(cext - cext_ref) <= 1e-5 * abs(cext_ref)
    h = s.history()
    assert len(h) == info["refinements"] + 1 and h[-1] < eps * eps


def test_mixed_c1_to_1e8(ctx):
    """Sphere 40 across, 80^3, to 1e-8 in mixed precision: several refinement steps."""
    g, params, cc, cs, prop, pol = cases.build("sphere_c1")
    einc = P.einc_plane(params, prop, pol)
    op = _op(ctx, g, params, cs)
    p64 = D.GpuQmrCs(op, max_iter=3000).get_pol_vec(einc, 1e-10)
    s = D.GpuQmrCs(op, max_iter=3000, mixed=True)
    p = s.get_pol_vec(einc, 1e-8)
    assert s.info["true_rel_residual"] < 1e-8 and s.info["refinements"] >= 2
    assert abs(P.ext_cross(params, p, einc) / P.ext_cross(params, p64, einc) - 1) < 1e-7
    assert abs(P.abs_cross(params, p, cc) / P.abs_cross(params, p64, cc) - 1) < 1e-7


@pytest.mark.parametrize("kind", ["bicgstab", "qmr"])
def test_mixed_solve_on_1024_point_columns(ctx, kind):
    """The largest-grid configuration's kernels at a small size: a box of
    8 x 8 x 500 dipoles (m = 1.6+0.05i, |m| kd = 0.5) on a 32 x 32 x 1024
    grid, so every Z pass runs on 1024-point columns (FP32: the 32 x 32 plan).
    Mixed solves to 1e-6 reach their true residual, and their Cext and Cabs
    agree with an FP64 QMR-CS solve to 1e-10 within 1e-5."""
    print("mixed 1024-point columns", kind, flush=True)
    m, d = 1.6 + 0.05j, 1e-7
    kd = 0.5 / abs(m)
    g = P.box(8, 8, 500, m, d)
    params = P.DDAParams(g, 2 * np.pi * d / kd, (32, 32, 1024))
    assert tuple(params.grid) == (32, 32, 1024)
    prop, pol1, _ = P.beam_pols((0, 0, 1))
    cc, cs = P.couple_constants(params, prop, pol1)
    einc = P.einc_plane(params, prop, pol1)
    op = _op(ctx, g, params, cs)
    p64 = D.GpuQmrCs(op, max_iter=20000).get_pol_vec(einc, 1e-10)
    s = D.SOLVER_CLASSES[kind](op, max_iter=20000, mixed=True)
    p = s.get_pol_vec(einc, 1e-6)
    assert s.info["true_rel_residual"] < 1e-6 and s.info["refinements"] >= 1
    assert abs(P.ext_cross(params, p, einc) / P.ext_cross(params, p64, einc) - 1) < 1e-5
    assert abs(P.abs_cross(params, p, cc) / P.abs_cross(params, p64, cc) - 1) < 1e-5


def test_mixed_c2_block_cext(ctx):
    """Sphere 128 dipoles across, 256^3 grid: both polarisations as one 2-RHS
    block, mixed BiCGStab to 1e-6; Cext and Cabs within 1e-5 of FP64 QMR-CS
    solved to 1e-8 on the GPU and of the CPU oracle's FP64 QMR-CS solve to
    1e-8 (tests/golden/c2_sphere128_fp64.json, made by make_c2_fixture.py);
    the GPU's FP64 solve within 1e-7 of the oracle's."""
    g, params, cc1, cs1, prop, pol1 = cases.build("sphere_c2")
    _, _, cc2, cs2, _, pol2 = cases.build("sphere_c2", pol_index=2)
    einc = np.array([P.einc_plane(params, prop, pol1), P.einc_plane(params, prop, pol2)])
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=64, grid=params.grid)
    op.set_couple_constants_batch(np.array([cs1, cs2]))
    print("c2 fp64 qmr", flush=True)
    p64, i64 = D.GpuQmrCs(op).get_pol_vecs(einc, 1e-8)
    print("c2 mixed bicgstab", [i["iterations"] for i in i64], flush=True)
    p, infos = D.GpuBicgStab(op, mixed=True).get_pol_vecs(einc, 1e-6)
    print([(i["iterations"], i["true_rel_residual"]) for i in infos], flush=True)
    import json
    import os
    with open(os.path.join(cases.GOLDEN, "c2_sphere128_fp64.json")) as f:
        fixture = json.load(f)
    assert fixture["nv"] == params.nv and tuple(fixture["grid"]) == tuple(params.grid)
    for r, cc in enumerate((cc1, cc2)):
        assert i64[r]["status"] == 0 and infos[r]["status"] == 0
        ref = fixture["pols"][r]
        c_ref, a_ref = ref["cext"], ref["cabs"]
        c64, a64 = P.ext_cross(params, p64[r], einc[r]), P.abs_cross(params, p64[r], cc)
        cm, am = P.ext_cross(params, p[r], einc[r]), P.abs_cross(params, p[r], cc)
        print("c2 pol %d vs CPU oracle: fp64 qmr cext %.2e cabs %.2e; mixed cext %.2e cabs %.2e" % (
            r + 1, abs(c64 / c_ref - 1), abs(a64 / a_ref - 1), abs(cm / c_ref - 1), abs(am / a_ref - 1)), flush=True)
        assert abs(c64 - c_ref) <= 1e-7 * abs(c_ref) and abs(a64 - a_ref) <= 1e-7 * abs(a_ref)
        assert abs(cm - c_ref) <= 1e-5 * abs(c_ref) and abs(am - a_ref) <= 1e-5 * abs(a_ref)
        assert infos[r]["true_rel_residual"] < 1e-6 and infos[r]["refinements"] >= 1
        c, c64 = P.ext_cross(params, p[r], einc[r]), P.ext_cross(params, p64[r], einc[r])
        assert abs(c - c64) <= 1e-5 * abs(c64)
        a, a64 = P.abs_cross(params, p[r], cc), P.abs_cross(params, p64[r], cc)
        assert abs(a - a64) <= 1e-5 * abs(a64)
        # Mie theory for the volume-equivalent sphere: the discretisation error
        # at |m| kd = 0.5 and 128 dipoles across is ~0.1%; the bound is 0.5%
        import math
        import oracle as O
        a_eq = (3 * params.nv / (4 * math.pi)) ** (1 / 3) * g.grid_unit
        qext, qsca, _, _ = O.mie(params.wave_num * a_eq, 1.5 + 0.1j)
        area = math.pi * a_eq ** 2
        assert abs(c / (qext * area) - 1) < 5e-3
        assert abs(a / ((qext - qsca) * area) - 1) < 5e-3


@pytest.mark.parametrize("kind", ["qmr", "bicg"])
def test_mixed_near_breakdown_rhs(ctx, kind):
    """An incident wave with |b^T b| / |b|^2 ~ 6e-6 (ellipsoid 32x24x16, the
    18th seeded direction): the complex-symmetric Lanczos start vector
    b / sqrt(b^T b) amplifies FP32 rounding ~1e5-fold and an FP32 inner solve
    stalls, so that step runs on the FP64 operator; the answer matches the
    FP64 solve."""
    import math
    m, d = 1.33 + 0.01j, 1e-7
    g = P.ellipsoid(32, 24, 16, m, d)
    params = P.DDAParams(g, 2 * math.pi * d * abs(m) / 0.6)
    einc, ccs, props, pols = P.incident_batch(params, P.directions(64, 2024)[18:19])
    b = einc[0]
    assert abs(np.sum(b * b)) / np.sum(abs(b) ** 2) < 1e-5
    op = _op(ctx, g, params, ccs[0])
    cls = D.SOLVER_CLASSES[kind]
    p64 = cls(op, max_iter=3000).get_pol_vec(b, 1e-8)
    s = cls(op, max_iter=3000, mixed=True)
    p = s.get_pol_vec(b, 1e-6)
    print(kind, s.info, flush=True)
    assert s.info["true_rel_residual"] < 1e-6 and s.info["iterations"] < 1000
    c, c64 = P.ext_cross(params, p, b), P.ext_cross(params, p64, b)
    assert abs(c - c64) <= 1e-5 * abs(c64)


def test_inner_plateau_stop_restarts_from_fp64_residual(ctx):
    """The mixed mode's plateau stop, forced from the host with a window of 8
    iterations and a ratio no solve can meet: every FP32 inner solve ends after
    ~8 iterations with its correction applied, and the outer FP64 steps
    restart it until the solve meets eps; the answer matches the FP64 solve.
    Invalid settings are refused."""
    g, params, cc, cs, prop, pol = cases.build("sphere_r7")
    einc = P.einc_plane(params, prop, pol)
    op = _op(ctx, g, params, cs)
    p64 = D.GpuQmrCs(op, max_iter=3000).get_pol_vec(einc, 1e-10)
    free = D.GpuQmrCs(op, max_iter=3000, mixed=True)
    p_free = free.get_pol_vec(einc, 1e-8)
    forced = D.GpuQmrCs(op, max_iter=3000, mixed=True, inner_plateau=(8, 1e-30))
    p = forced.get_pol_vec(einc, 1e-8)
    print("free", free.info, "forced", forced.info, flush=True)
    assert forced.info["true_rel_residual"] < 1e-8
    assert forced.info["refinements"] > free.info["refinements"]
    assert cases.rel_err(p, p64) <= 1e-6 and cases.rel_err(p_free, p64) <= 1e-6
    with pytest.raises(D.DDAError):
        D.GpuQmrCs(op, inner_plateau=(8, 0.5))  # not a mixed-mode solver
    with pytest.raises(D.DDAError):
        D.GpuQmrCs(op, mixed=True, inner_plateau=(8, 0.0))
