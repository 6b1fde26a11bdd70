"""The allocation-free cross sections (dda_cross_sections_noalloc) against the
allocating entry point (dda_cross_sections), both called through the C ABI:
two calls in a row, the second with fewer dipoles and right-hand sides, on
random polarisations and on a solved two-material sphere; and the size limit
of the fixed scratch array, which is rejected on the host."""
import ctypes

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _both(op, p_dev, e_dev, nrhs, cc):
    """(cext, cabs) of the allocating and of the allocation-free entry point."""
    L = D.lib()
    ccf = np.ascontiguousarray(np.asarray(cc, np.complex128).reshape(nrhs, -1, 3)).view(np.float64)
    out = []
    for fn in (L.dda_cross_sections, L.dda_cross_sections_noalloc):
        cext, cabs = np.full(nrhs, np.nan), np.full(nrhs, np.nan)
        rc = fn(op.handle, p_dev.ptr, e_dev.ptr, int(nrhs), _ptr(ccf), _ptr(cext), _ptr(cabs))
        assert rc == D.OK, D.lib().dda_last_error()
        out.append((cext, cabs))
    return out


def _random_case(ctx, radius, ndir, seed):
    g = P.sphere(radius, 1.5 + 0.1j, 0.5e-6)
    params = P.DDAParams(g, 6e-6)
    dirs = P.directions(ndir, seed)
    einc, ccs, props, pols = P.incident_batch(params, dirs)
    k = len(ccs)
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid)
    rng = np.random.default_rng(seed)
    p = np.zeros((k, 3, params.vec_stride), np.complex128)
    p[:, :, :params.nv] = rng.uniform(-1, 1, (k, 3, params.nv)) + 1j * rng.uniform(-1, 1, (k, 3, params.nv))
    p_dev = D.DeviceVector(ctx, k * params.vec_size, 64)
    p_dev.upload(p.reshape(-1))
    e_dev = D.DeviceVector(ctx, k * params.vec_size, 64)
    e_dev.upload(np.asarray(einc).reshape(-1))
    cc = np.array([P.couple_constants(params, props[r], pols[r])[0] for r in range(k)])
    return op, p_dev, e_dev, k, cc


def _check(res):
    (ce, ca), (ce0, ca0) = res
    assert np.all(np.isfinite(ce0)) and np.all(np.isfinite(ca0))
    # the same partial sums in the same order: Cext to the last bit; Cabs forms
    # -Im(1/alpha) on the device instead of the host
    np.testing.assert_allclose(ce0, ce, rtol=1e-15, atol=0)
    np.testing.assert_allclose(ca0, ca, rtol=1e-13, atol=0)


def test_noalloc_matches_allocating_entry_point_twice(ctx):
    big = _random_case(ctx, 12.0, 3, 11)      # 24 across, 6 right-hand sides
    small = _random_case(ctx, 5.0, 1, 12)     # 10 across, 2 right-hand sides
    _check(_both(*big))
    _check(_both(*small))                     # the second call: smaller input, same scratch
    _check(_both(*big))
    for op, p_dev, e_dev, _, _ in (big, small):
        for v in (p_dev, e_dev):
            v.close()
        op.close()


def test_noalloc_on_solved_two_material_sphere(ctx):
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    einc, ccs, props, pols = P.incident_batch(params, np.array([(0, 0, 1), (0.6, 0, 0.8)], np.float64))
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid,
                     nmat=len(g.materials))
    op.set_couple_constants_batch(ccs)
    k = len(ccs)
    e_dev = D.incident_plane(op, props, pols)
    p_dev = D.DeviceVector(ctx, k * params.vec_size, 64)
    s = D.GpuQmrCs(op, max_iter=5000)
    infos = s.solve_device(e_dev, p_dev, k, 1e-10)
    assert all(i["status"] == D.OK for i in infos)
    ccf = np.array([P.couple_constants(params, props[r], pols[r])[0] for r in range(k)])
    _check(_both(op, p_dev, e_dev, k, ccf))
    # Cext only (no polarisabilities): the second sum is not asked for
    L = D.lib()
    cext = np.zeros(k)
    assert L.dda_cross_sections_noalloc(op.handle, p_dev.ptr, e_dev.ptr, k, None, _ptr(cext), None) == D.OK
    ref, _ = D.cross_sections(op, p_dev, e_dev, k)
    np.testing.assert_allclose(cext, ref, rtol=1e-15, atol=0)
    s.close()
    for v in (p_dev, e_dev):
        v.close()
    op.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(ctx):
    g = P.sphere(3.0, 1.5 + 0.1j, 0.5e-6)
    params = P.DDAParams(g, 6e-6)
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid)
    # 1 block per right-hand side: 6 + 4 doubles each; 2^20 / 10 right-hand sides fit
    k = (1 << 20) // 10 + 1
    L = D.lib()
    cc = np.ones((k, 1, 3), np.complex128).view(np.float64)
    cext, cabs = np.zeros(k), np.zeros(k)
    p_dev = D.DeviceVector(ctx, 16, 64)  # never read: the call is refused before any launch
    rc = L.dda_cross_sections_noalloc(op.handle, p_dev.ptr, p_dev.ptr, k, _ptr(cc), _ptr(cext), _ptr(cabs))
    assert rc == D.INVALID_ARG
    assert b"scratch" in L.dda_last_error()
    p_dev.close()
    op.close()
