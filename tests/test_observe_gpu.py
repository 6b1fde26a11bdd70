"""Device-side steps around the solve against the oracle: incident field,
polarisabilities, device-resident solve, Cext/Cabs, far fields, Jones matrices."""
import math

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu


def _solved(ctx, name="two_materials", dirs=((0, 0, 1),)):
    g, params, cc, cs, prop, pol = cases.build(name)
    einc, ccs, props, pols = P.incident_batch(params, np.array(dirs, np.float64))
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid,
                     nmat=len(g.materials))
    op.set_couple_constants_batch(ccs)
    k = len(ccs)
    e_dev = D.incident_plane(op, props, pols)
    p_dev = D.DeviceVector(ctx, k * params.vec_size, 64)
    s = D.GpuQmrCs(op, max_iter=5000)
    infos = s.solve_device(e_dev, p_dev, k, 1e-10)
    assert all(i["status"] == D.OK for i in infos)
    return g, params, op, s, einc, ccs, props, pols, e_dev, p_dev


def test_incident_and_device_solve(ctx):
    g, params, op, s, einc, ccs, props, pols, e_dev, p_dev = _solved(ctx, dirs=((0, 0, 1), (0.6, 0, 0.8)))
    e = e_dev.download().reshape(len(ccs), -1)
    np.testing.assert_allclose(e, einc, rtol=0, atol=1e-14)
    # the device-resident solve is bitwise the host-array one on the same input
    p_host, _ = s.get_pol_vecs(e, 1e-10)
    assert np.array_equal(p_dev.download().reshape(len(ccs), -1), p_host)
    # and agrees with the solve of the host-built field to the solver tolerance
    p_host2, _ = s.get_pol_vecs(einc, 1e-10)
    assert cases.rel_err(p_host, p_host2) <= 1e-8


def test_cross_sections_match_oracle(ctx):
    g, params, op, s, einc, ccs, props, pols, e_dev, p_dev = _solved(ctx)
    k = len(ccs)
    ccf = np.array([P.couple_constants(params, props[r], pols[r])[0] for r in range(k)])
    cext, cabs = D.cross_sections(op, p_dev, e_dev, k, ccf)
    p = p_dev.download().reshape(k, -1)
    for r in range(k):
        orc = cases.oracle_problem(g, params, ccs[r])
        assert abs(cext[r] - orc.ext_cross_sum(p[r], einc[r])) <= 1e-12 * abs(cext[r])
        assert abs(cabs[r] - orc.abs_cross(p[r], ccf[r])) <= 1e-12 * abs(cabs[r])


def test_far_field_and_jones_match_oracle(ctx):
    g, params, op, s, einc, ccs, props, pols, e_dev, p_dev = _solved(ctx)
    k = len(ccs)
    p = p_dev.download().reshape(k, -1)
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(37, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    rhs = rng.integers(0, k, 37)
    ff = D.far_field(op, p_dev, k, dirs, rhs)
    orc = cases.oracle_problem(g, params, ccs[0])
    for i in range(37):
        ref = orc.calc_field(p[rhs[i]], dirs[i])
        assert np.linalg.norm(ff[i] - ref) <= 1e-12 * np.linalg.norm(ref)
    # forward amplitude: optical theorem against the Cext sum
    fwd = orc.ext_cross_forward(p[0], props[0], pols[0])
    assert abs(fwd - orc.ext_cross_sum(p[0], einc[0])) <= 1e-6 * abs(fwd)
    args = (props[0], pols[0], pols[1], 7, 0.0, math.radians(30), 5, 0.0, math.radians(72))
    J = D.jones(op, p_dev, 0, *args)
    Jr = orc.jones(p[0], p[1], 0, *args)
    assert np.max(np.abs(J - Jr)) <= 1e-12 * np.max(np.abs(Jr))
    Js = D.jones(op, p_dev, 1, *args)
    Jsr = orc.jones(p[0], None, 1, *args)
    assert np.max(np.abs(Js - Jsr)) <= 1e-12 * np.max(np.abs(Jsr))


def test_symmetric_jones_on_chiral_target(ctx):
    """FarFieldCalc.cpp:163-168 on the device: for the chiral pinwheel (symmetric
    under the quarter turn about z that takes polarisation 1 to 2, and under no
    mirror plane), the Jones matrices from the polarisation-1 solve alone
    (symmetric = 1) equal those from both solves to the solver's tolerance, and
    equal the oracle's symmetric evaluation of the same solution."""
    g, params, op, s, einc, ccs, props, pols, e_dev, p_dev = _solved(ctx, "pinwheel")
    args = (props[0], pols[0], pols[1], 13, 0.0, math.pi / 12, 12, 0.0, math.pi / 6)
    J2 = D.jones(op, p_dev, 0, *args)
    J1 = D.jones(op, p_dev, 1, *args)
    scale = np.max(np.abs(J2))
    assert np.max(np.abs(J1 - J2)) <= 1e-8 * scale
    p = p_dev.download().reshape(len(ccs), -1)
    orc = cases.oracle_problem(g, params, ccs[0])
    Jr = orc.jones(p[0], None, 1, *args)
    assert np.max(np.abs(J1 - Jr)) <= 1e-12 * np.max(np.abs(Jr))


def test_anisotropic_material_solve_and_hdf5_output(ctx, tmp_path):
    """An anisotropic (DiagMatrix3) material end to end: polarisabilities from
    dda_couple_constants_diag (each axis its own index,
    PolarizabilityDescription.cpp:69-75, 90-96), a device QMR-CS solve against
    the oracle's, and the solution written as the reference's DipPol HDF5 file
    and read back bit for bit."""
    from dda_amd import hdf5 as H
    g, params, _, _, prop, pol = cases.build("sphere_r4")
    g.materials = [np.array([1.5 + 0.1j, 1.6 + 0.05j, 1.4 + 0.2j])]
    cc, cs = D.couple_constants(g.materials, g.grid_unit, params.kd, prop, pol, "ldr")
    assert abs(cs[0, 0] - cs[0, 1]) > 1e-2 * abs(cs[0, 0])  # the axes differ
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid)
    op.set_couple_constants(cs)
    einc = P.einc_plane(params, prop, pol)
    s = D.GpuQmrCs(op, max_iter=2000)
    p = s.get_pol_vec(einc, 1e-10)
    orc = cases.oracle_problem(g, params, cs)
    st, p_ref, _, _ = orc.solve("qmr", einc, 1e-10, max_iter=2000)
    assert st == 0 and cases.rel_err(p, p_ref) <= 1e-8
    assert abs(P.abs_cross(params, p, cc) / orc.abs_cross(p_ref, cc) - 1) <= 1e-8
    assert not P.rotation_symmetric(g, prop)  # anisotropic: both polarisations are solved
    if H.available():
        f = tmp_path / "DipPol-Pol1.hdf5"
        _, p1, p2 = P.beam_pols(prop)
        H.write_dippol(f, params, p, 1, prop, [p1, p2])
        assert np.array_equal(H.read_dippol(f, params), p)
        mats = H.read_dataset(f, "/Geometry/RefractiveIndices", complex_=True)
        np.testing.assert_array_equal(mats[0], g.materials[0])
    s.close()
    op.close()
