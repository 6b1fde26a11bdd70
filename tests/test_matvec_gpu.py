"""Parity of the HIP matvec with the oracle (FP64 <= 1e-12, FP32 <= 1e-5 relative).

The oracle here is the long-double CPU restatement of MatVecCpu
(DDA/MatVecCpu.cpp:67-150) with the D matrix of DDA/DMatrixCpu.cpp:133-205.
"""
import os

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 1e-5


def _ops(ctx, g, params, cs, precision, **kw):
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=precision,
                     grid=params.grid, nmat=len(g.materials), **kw)
    op.set_couple_constants(cs)
    return op


@pytest.mark.parametrize("name", cases.SMALL_CASES)
@pytest.mark.parametrize("conj", [False, True])
def test_matvec_fp64_matches_oracle(ctx, name, conj):
    print("case", name, "conj", conj, flush=True)
    g, params, cc, cs, prop, pol = cases.build(name)
    orc = cases.oracle_problem(g, params, cs)
    x = P.random_dipole_vector(params, seed=7)
    y_ref = orc.matvec(x, conj=conj, precision=80)
    op = _ops(ctx, g, params, cs, 64)
    y = D.MatVecGpu(op).apply(x, conj=conj)
    assert cases.rel_err(y, y_ref) <= FP64_TOL
    assert np.all(y[params.nv:params.vec_stride] == 0)


@pytest.mark.parametrize("name", cases.SMALL_CASES)
def test_matvec_fp32_matches_oracle(ctx, name):
    print("case", name, flush=True)
    g, params, cc, cs, prop, pol = cases.build(name)
    orc = cases.oracle_problem(g, params, cs)
    x = P.random_dipole_vector(params, seed=8)
    y_ref = orc.matvec(x, precision=80)
    op = _ops(ctx, g, params, cs, 32)
    y = D.MatVecGpu(op).apply(x)
    assert cases.rel_err(y, y_ref) <= FP32_TOL


@pytest.mark.parametrize("grid,pot", [((0, 0, 0), False), ((24, 20, 16), True), ((30, 18, 14), True)])
def test_matvec_grid_choices(ctx, grid, pot):
    """Power-of-two, larger and smooth grids give the same operator (exact embedding)."""
    print("grid", grid, pot, flush=True)
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=grid, support_non_pot=pot)
    orc = cases.oracle_problem(g, params, cs)
    x = P.random_dipole_vector(params, seed=9)
    y_ref = orc.matvec(x, precision=80)
    y = D.MatVecGpu(_ops(ctx, g, params, cs, 64)).apply(x)
    assert cases.rel_err(y, y_ref) <= FP64_TOL


# Grid lengths with compile-time planned kernels, each on the Y (passes 2, 4),
# the Z (pass 3) and the X axis (passes 1, 5); the 10x8x6 ellipsoid box embeds
# exactly in all.
STATIC_GRIDS = [(32, 64, 96), (32, 96, 64), (32, 128, 192), (32, 192, 128), (32, 256, 512), (32, 512, 256),
                (32, 1024, 16), (32, 16, 1024),
                # and on the X axis (passes 1, 5)
                (64, 16, 16), (96, 16, 16), (128, 16, 16), (192, 16, 16), (256, 16, 16), (512, 16, 16),
                (1024, 16, 16), (32, 32, 32), (48, 48, 48)]


@pytest.mark.parametrize("grid", STATIC_GRIDS)
@pytest.mark.parametrize("precision", [64, 32])
def test_matvec_static_plans(ctx, grid, precision):
    print("grid", grid, "precision", precision, flush=True)
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=grid)
    assert params.grid == grid
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=11)
    y_ref = orc.matvec(x, precision=64)
    y = D.MatVecGpu(_ops(ctx, g, params, cs, precision)).apply(x)
    assert cases.rel_err(y, y_ref) <= (FP64_TOL if precision == 64 else FP32_TOL)


@pytest.mark.parametrize("grid", [(32, 16, 1024), (48, 24, 1024)])
@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_matvec_z1024_variants(ctx, grid, variant):
    """Every Z-pass variant at 1024 points (narrow and wide plans; the 32 x 32
    plan over 16 and over 8 columns per work-group), pinned through
    dda_op_set_variants on an FP32 operator, against the oracle; 48 columns
    leave a partial tile for 32 and for 16 columns per work-group."""
    print("grid", grid, "variant", variant, flush=True)
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=grid)
    assert params.grid == grid
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=12)
    y_ref = orc.matvec(x, precision=64)
    op = _ops(ctx, g, params, cs, 32)
    assert op.variants([-1, variant, -1, -1, -1, -1])[1] == variant
    for conj in (False, True):
        y = D.MatVecGpu(op).apply(x, conj=conj)
        y_ref = orc.matvec(x, conj=conj, precision=64)
        assert cases.rel_err(y, y_ref) <= FP32_TOL
    op.close()


@pytest.mark.parametrize("name,grid", [("sphere_r7", (32, 32, 32)), ("ellipsoid_5_4_3", (32, 16, 16)),
                                       ("sphere_r7", (0, 0, 0)), ("cylinder_9_3", (0, 0, 0)),
                                       ("ellipsoid_5_4_3", (30, 18, 14))])
@pytest.mark.parametrize("precision", [64, 32])
def test_rows_without_dipoles_stay_zero(ctx, name, grid, precision):
    """The grid buffers keep nothing of an earlier apply: a second apply after
    a large first one still matches the oracle, rows of the (z, y) plane
    without dipoles included (compile-time planned passes on the 32-point
    grids, run-time planned ones on 28^3, 18 x 12 x 12 and 30 x 18 x 14)."""
    print("case", name, grid, precision, flush=True)
    g, params, cc, cs, prop, pol = cases.build(name, grid=grid, support_non_pot=True)
    orc = cases.oracle_problem(g, params, cs)
    mv = D.MatVecGpu(_ops(ctx, g, params, cs, precision))
    mv.apply(1e6 * P.random_dipole_vector(params, seed=21))
    x = P.random_dipole_vector(params, seed=22)
    for conj in (False, True):
        y = mv.apply(x, conj=conj)
        assert cases.rel_err(y, orc.matvec(x, conj=conj, precision=80)) <= (FP64_TOL if precision == 64 else FP32_TOL)


@pytest.mark.parametrize("grid", [(32, 64, 96), (32, 96, 64), (32, 128, 192), (32, 192, 128), (32, 256, 512),
                                  (32, 512, 256), (32, 1024, 16), (32, 16, 1024), (32, 48, 32), (96, 16, 16),
                                  (256, 16, 16), (1024, 16, 16)])
@pytest.mark.parametrize("variant", [2, 3])
def test_matvec_other_tile_variants(ctx, grid, variant):
    """Every FP32 pass with the other tile (more or fewer columns or lines per
    work-group; at 1024 Z points the 32 x 32 plans), narrow (2) and wide (3),
    pinned through dda_op_set_variants, against the oracle."""
    print("grid", grid, "variant", variant, flush=True)
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=grid)
    assert params.grid == grid
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=14)
    op = _ops(ctx, g, params, cs, 32)
    assert op.variants([variant] * 5 + [-1])[:5] == [variant] * 5
    for conj in (False, True):
        y = D.MatVecGpu(op).apply(x, conj=conj)
        assert cases.rel_err(y, orc.matvec(x, conj=conj, precision=64)) <= FP32_TOL
    op.close()


@pytest.mark.parametrize("grid", [(32, 64, 96), (32, 96, 64), (32, 128, 192), (32, 192, 128), (32, 256, 512),
                                  (32, 512, 256), (32, 1024, 16), (32, 48, 32)])
@pytest.mark.parametrize("variant", [4, 5])
def test_matvec_third_tile_y_variants(ctx, grid, variant):
    """The FP32 Y passes with the third tile (four times the default columns
    where they fit), narrow (4) and wide (5), against the oracle."""
    print("grid", grid, "variant", variant, flush=True)
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=grid)
    assert params.grid == grid
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=15)
    op = _ops(ctx, g, params, cs, 32)
    assert op.variants([variant, -1, variant, -1, -1, -1])[0] == variant
    for conj in (False, True):
        y = D.MatVecGpu(op).apply(x, conj=conj)
        assert cases.rel_err(y, orc.matvec(x, conj=conj, precision=64)) <= FP32_TOL
    op.close()


def test_variants_rejects_out_of_range(ctx):
    g, params, cc, cs, prop, pol = cases.build("sphere_r4")
    op = _ops(ctx, g, params, cs, 32)
    with pytest.raises(D.DDAError):
        op.variants([-1, 4, -1, -1, -1, -1])
    with pytest.raises(D.DDAError):
        op.variants([6, -1, -1, -1, -1, -1])
    with pytest.raises(D.DDAError):
        op.variants([-1, -1, -1, 4, -1, -1])
    with pytest.raises(D.DDAError):
        op.variants([-1, -1, -1, -1, -1, 2])
    assert op.variants() == [-1, -1, -1, -1, -1, -1]
    op.close()


def test_dmatrix_device_generation_matches_oracle(ctx):
    """FP64 tensor generated on the GPU vs DMatrixCpu::createDMatrix restated."""
    import oracle as O
    g, params, cc, cs, prop, pol = cases.build("box_10_8_6")
    op = _ops(ctx, g, params, cs, 64)
    dm = op.dmatrix()
    ref = O.dmatrix(g.box, params.grid, g.grid_unit, params.wave_num)
    assert np.max(np.abs(dm - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_host_dmatrix_input(ctx):
    """A D matrix handed over in the reference layout gives the same operator."""
    import oracle as O
    g, params, cc, cs, prop, pol = cases.build("sphere_r4")
    ref = O.dmatrix(g.box, params.grid, g.grid_unit, params.wave_num)
    op = _ops(ctx, g, params, cs, 64, dmatrix=ref)
    x = P.random_dipole_vector(params, seed=10)
    y_ref = cases.oracle_problem(g, params, cs).matvec(x, precision=80)
    y = D.MatVecGpu(op).apply(x)
    assert cases.rel_err(y, y_ref) <= FP64_TOL


def test_matvec_c1_full_size(ctx):
    """Sphere 40 dipoles across (33,552 dipoles), 80^3 grid, FP64 and FP32."""
    g, params, cc, cs, prop, pol = cases.build("sphere_c1")
    assert params.nv == 33552 and params.grid == (80, 80, 80)
    orc = cases.oracle_problem(g, params, cs)
    x = P.random_dipole_vector(params)
    y_ref = orc.matvec(x, precision=64)
    y64 = D.MatVecGpu(_ops(ctx, g, params, cs, 64)).apply(x)
    assert cases.rel_err(y64, y_ref) <= FP64_TOL
    y32 = D.MatVecGpu(_ops(ctx, g, params, cs, 32)).apply(x)
    assert cases.rel_err(y32, y_ref) <= FP32_TOL


def test_matvec_c2_full_size(ctx):
    """Sphere 128 dipoles across (1,099,136 dipoles), 256^3 grid: FP64 (both
    conj flags) and FP32 against the oracle with its own FP64 tensor."""
    import oracle as O
    g, params, cc, cs, prop, pol = cases.build("sphere_c2")
    assert params.nv == 1099136 and params.grid == (256, 256, 256)
    O.set_threads(min(16, os.cpu_count() or 1))
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params)
    mv64 = D.MatVecGpu(_ops(ctx, g, params, cs, 64))
    y_ref = orc.matvec(x)
    print("c2 fp64", flush=True)
    assert cases.rel_err(mv64.apply(x), y_ref) <= FP64_TOL
    print("c2 fp64 conj", flush=True)
    assert cases.rel_err(mv64.apply(x, conj=True), orc.matvec(x, conj=True)) <= FP64_TOL
    print("c2 fp32", flush=True)
    y32 = D.MatVecGpu(_ops(ctx, g, params, cs, 32)).apply(x)
    assert cases.rel_err(y32, y_ref) <= FP32_TOL


def test_matvec_c5_full_size(ctx):
    """Sphere 512 dipoles across (70,277,000 dipoles), 1024^3 grid, device
    vectors: the FP32 matvec within 1e-5 of the FP64 one (the 1024-point
    kernels of both precisions against each other; the oracle is too slow at
    this size), padding zero."""
    import math
    m, d = 1.6 + 0.05j, 1e-7
    g = P.sphere(256.0, m, d)
    params = P.DDAParams(g, 2 * math.pi * d * abs(m) / 0.5)
    assert params.nv == 70277000 and params.grid == (1024, 1024, 1024)
    prop, pol1, _ = P.beam_pols((0, 0, 1))
    cc, cs = P.couple_constants(params, prop, pol1)
    x = P.random_dipole_vector(params)
    ys = {}
    for prec in (64, 32):
        print("c5 precision", prec, flush=True)
        op = _ops(ctx, g, params, cs, prec)
        xd, yd = D.DeviceVector(ctx, params.vec_size, prec), D.DeviceVector(ctx, params.vec_size, prec)
        xd.upload(x)
        op.apply(xd, yd)
        ys[prec] = yd.download()
        for o in (op, xd, yd):
            o.close()
    assert np.all(ys[32][params.nv:params.vec_stride] == 0)
    assert cases.rel_err(ys[32].astype(np.complex128), ys[64]) <= FP32_TOL


def test_matvec_complex_symmetry_large(ctx):
    """Size-independent property at a 128^3 grid: x^T (A y) = y^T (A x)."""
    g = P.sphere(32.0, 1.5 + 0.1j, 0.5e-6)
    params = P.DDAParams(g, 6e-6)
    prop, pol1, _ = P.beam_pols((0, 0, 1))
    cc, cs = P.couple_constants(params, prop, pol1)
    mv = D.MatVecGpu(_ops(ctx, g, params, cs, 64))
    x = P.random_dipole_vector(params, seed=1)
    y = P.random_dipole_vector(params, seed=2)
    a = np.dot(x, mv.apply(y))
    b = np.dot(y, mv.apply(x))
    assert abs(a - b) <= 1e-12 * abs(a)
    # A^H consistency: <A x, y> = <x, A^H y>
    c = np.vdot(mv.apply(x), y)
    d = np.vdot(x, mv.apply(y, conj=True))
    assert abs(c - d) <= 1e-12 * abs(c)


def test_device_apply_matches_host_apply(ctx):
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    op = _ops(ctx, g, params, cs, 64)
    x = P.random_dipole_vector(params, seed=3)
    xd = D.DeviceVector(ctx, params.vec_size, 64)
    yd = D.DeviceVector(ctx, params.vec_size, 64)
    xd.upload(x)
    op.apply(xd, yd)
    ctx.synchronize()
    np.testing.assert_array_equal(yd.download(), D.MatVecGpu(op).apply(x))


def test_invalid_arguments(ctx):
    g, params, cc, cs, prop, pol = cases.build("sphere_r4")
    with pytest.raises(D.DDAError):
        D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=(8, 8, 8))  # grid < 2*box
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num)
    with pytest.raises(D.DDAError):
        D.MatVecGpu(op).apply(np.zeros(params.vec_size, complex))  # couple constants not set
