"""Batched right-hand sides: block matvec and lockstep Krylov solves.

Each right-hand side of a batch must give bitwise the result it gets alone
(same kernels, per-RHS reductions whose shape depends only on the vector
length), and the single-RHS results are pinned to the oracle elsewhere; one
RHS per case is checked against the oracle here too.
"""
import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu


def _setup(name, ndirs, seed, grid=(0, 0, 0)):
    g, params, cc, cs, prop, pol = cases.build(name, grid=grid)
    einc, ccs, props, pols = P.incident_batch(params, P.directions(ndirs, seed))
    return g, params, einc, ccs


def _op(ctx, g, params, precision=64):
    return D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=precision,
                       grid=params.grid, nmat=len(g.materials))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("conj", [False, True])
@pytest.mark.parametrize("name,grid", [("two_materials", (0, 0, 0)), ("ellipsoid_5_4_3", (32, 64, 96))])
def test_block_matvec_bitwise(ctx, precision, conj, name, grid):
    print("block matvec", name, grid, precision, conj, flush=True)
    g, params, einc, ccs = _setup(name, 3, 11, grid)
    k, n = len(ccs), params.vec_size
    rng = np.random.default_rng(5)
    x = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex128)
    for r in range(k):
        x[r, params.nv:params.vec_stride] = 0
    op = _op(ctx, g, params, precision)
    xs, ys = D.DeviceVector(ctx, n, precision), D.DeviceVector(ctx, n, precision)
    alone = []
    for r in range(k):
        op.set_couple_constants(ccs[r])
        xs.upload(x[r])
        op.apply(xs, ys, conj)
        alone.append(ys.download())
    op.set_couple_constants_batch(ccs)
    xb, yb = D.DeviceVector(ctx, k * n, precision), D.DeviceVector(ctx, k * n, precision)
    xb.upload(x.reshape(-1))
    op.apply_batch(xb, yb, k, conj)
    yb = yb.download().reshape(k, n)
    for r in range(k):
        assert np.array_equal(yb[r], alone[r]), r
    # one RHS against the oracle at the precision's bar
    orc = cases.oracle_problem(g, params, ccs[k - 1])
    y_ref = orc.matvec(x[k - 1], conj=conj, precision=80)
    assert cases.rel_err(yb[k - 1].astype(np.complex128), y_ref) <= (1e-12 if precision == 64 else 1e-5)


def test_block_matvec_needs_cc_for_every_rhs(ctx):
    g, params, einc, ccs = _setup("sphere_r4", 2, 3)
    op = _op(ctx, g, params)
    op.set_couple_constants(ccs[0])
    n = params.vec_size
    xb, yb = D.DeviceVector(ctx, 4 * n, 64), D.DeviceVector(ctx, 4 * n, 64)
    with pytest.raises(D.DDAError) as e:
        op.apply_batch(xb, yb, 4)
    assert e.value.status == D.INVALID_ARG


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("kind", ["qmr", "bicgstab", "cgnr", "bicg"])
def test_batched_solve_bitwise(ctx, kind, mixed):
    print("batched solve", kind, "mixed", mixed, flush=True)
    g, params, einc, ccs = _setup("ellipsoid_5_4_3", 2, 17)
    eps = 1e-8
    op = _op(ctx, g, params)
    s = D.SOLVER_CLASSES[kind](op, max_iter=5000, mixed=mixed)
    alone, infos = [], []
    for r in range(len(ccs)):
        op.set_couple_constants(ccs[r])
        alone.append(s.get_pol_vec(einc[r], eps))
        infos.append(dict(s.info))
    op.set_couple_constants_batch(ccs)
    pb, binfo = s.get_pol_vecs(einc, eps)
    for r in range(len(ccs)):
        assert binfo[r]["status"] == D.OK
        assert binfo[r]["iterations"] == infos[r]["iterations"], r
        assert np.array_equal(pb[r], alone[r]), r
        assert binfo[r]["true_rel_residual"] <= 10 * eps
    # RHS 1 against the oracle's FP64 solve
    orc = cases.oracle_problem(g, params, ccs[1])
    st, p_ref, it_ref, _ = orc.solve(kind, einc[1], eps, max_iter=5000)
    assert st == 0
    assert cases.rel_err(pb[1], p_ref) <= 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_batch_zero_rhs_and_shared_cc(ctx, mixed):
    """A zero field fails alone (INVALID_ARG, zero output); the others are
    unaffected. One set of couple constants serves every RHS."""
    g, params, einc, ccs = _setup("sphere_r4", 1, 2)
    op = _op(ctx, g, params)
    op.set_couple_constants(ccs[0])
    s = D.GpuQmrCs(op, max_iter=2000, mixed=mixed)
    e0 = einc[0]
    ref = s.get_pol_vec(e0, 1e-9)
    batch = np.stack([np.zeros_like(e0), e0, 2 * e0])
    pb, info = s.get_pol_vecs(batch, 1e-9)
    assert info[0]["status"] == D.INVALID_ARG and not np.any(pb[0])
    assert info[1]["status"] == D.OK and np.array_equal(pb[1], ref)
    assert info[2]["status"] == D.OK
    assert cases.rel_err(pb[2], 2 * ref) <= 1e-8


def test_batch_max_iter_per_rhs(ctx):
    g, params, einc, ccs = _setup("sphere_r4", 1, 2)
    op = _op(ctx, g, params)
    op.set_couple_constants_batch(ccs)
    s = D.GpuQmrCs(op, max_iter=3)
    pb, info = s.get_pol_vecs(einc, 1e-13)
    assert [i["status"] for i in info] == [D.MAX_ITER, D.MAX_ITER]
    assert [i["iterations"] for i in info] == [3, 3]


def test_c3_ellipsoid_128_rhs(ctx):
    """Ellipsoid, semi-axes 32x24x16 dipoles, m = 1.33+0.01i, |m|kd = 0.6,
    64 seeded directions x 2 polarisations, mixed block QMR-CS to 1e-6."""
    import math
    m = 1.33 + 0.01j
    d = 1e-7
    g = P.ellipsoid(32, 24, 16, m, d)
    params = P.DDAParams(g, 2 * math.pi * d * abs(m) / 0.6)
    einc, ccs, props, pols = P.incident_batch(params, P.directions(64, 2024))
    assert einc.shape[0] == 128
    op = _op(ctx, g, params)
    op.set_couple_constants_batch(ccs)
    s = D.GpuQmrCs(op, mixed=True)
    pb, info = s.get_pol_vecs(einc, 1e-6)
    print("ellipsoid batch iterations", [i["iterations"] for i in info][:8], "seconds", info[0]["seconds"], flush=True)
    bad = [(r, i) for r, i in enumerate(info) if i["status"] != D.OK]
    assert not bad, bad[:4]
    assert max(i["true_rel_residual"] for i in info) < 1e-6
    # a few RHS: FP64 QMR-CS alone gives the same Cext to 1e-5
    s64 = D.GpuQmrCs(op)
    for r in (0, 77, 127):
        op.set_couple_constants(ccs[r])
        p64 = s64.get_pol_vec(einc[r], 1e-6)
        c_mixed = P.ext_cross(params, pb[r], einc[r])
        c_64 = P.ext_cross(params, p64, einc[r])
        assert abs(c_mixed - c_64) <= 1e-5 * abs(c_64), (r, c_mixed, c_64)
    # and RHS 77 is bitwise the mixed solve alone
    op.set_couple_constants(ccs[77])
    assert np.array_equal(s.get_pol_vec(einc[77], 1e-6), pb[77])


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_start_vectors(ctx, mixed):
    """Batched solves from start vectors (--load-start-dip-pol per field): each
    field's result is bitwise the one it gets alone from the same start."""
    g, params, einc, ccs = _setup("sphere_r4", 2, 9)
    op = _op(ctx, g, params)
    s = D.GpuBicgStab(op, max_iter=2000, mixed=mixed)
    op.set_couple_constants_batch(ccs)
    p0, _ = s.get_pol_vecs(einc, 1e-4)  # rough solutions as starts
    alone = []
    for r in range(len(ccs)):
        op.set_couple_constants(ccs[r])
        alone.append(s.get_pol_vec(einc[r], 1e-10, start=p0[r]))
        assert s.info["true_rel_residual"] <= 1e-9
    op.set_couple_constants_batch(ccs)
    pb, infos = s.get_pol_vecs(einc, 1e-10, start=p0)
    for r in range(len(ccs)):
        assert np.array_equal(pb[r], alone[r]), r


@pytest.mark.parametrize("precision,mode", [(64, "fp64"), (64, "mixed")])
def test_split_block_matches_one_stream(ctx, precision, mode):
    """A block of right-hand sides runs as two halves on two streams (the Z
    pass of one beside the X and Y passes of the other); the block matvec and
    a block solve (whose steady iterations replay a captured graph holding the
    fork and join) give bitwise what the block gives on one stream, over
    repeated solves."""
    print("split block", mode, flush=True)
    g, params, einc, ccs = _setup("two_materials", 3, 17)  # 6 right-hand sides
    k, n = len(ccs), params.vec_size
    op1 = _op(ctx, g, params, precision)
    op2 = _op(ctx, g, params, precision)
    assert op1.variants()[5] == -1  # off by default
    op2.variants([-1, -1, -1, -1, -1, 1])
    rng = np.random.default_rng(8)
    x = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex128)
    for r in range(k):
        x[r, params.nv:params.vec_stride] = 0
    xb = D.DeviceVector(ctx, k * n, precision)
    xb.upload(x.reshape(-1))
    outs = []
    for op in (op1, op2):
        op.set_couple_constants_batch(ccs)
        yb = D.DeviceVector(ctx, k * n, precision)
        op.apply_batch(xb, yb, k)
        outs.append(yb.download())
    assert np.array_equal(outs[0], outs[1])
    sols = []
    for op in (op1, op2):
        s = D.GpuBicgStab(op, max_iter=2000, mode=mode)
        runs = [s.get_pol_vecs(einc, 1e-9) for _ in range(3)]
        for p, infos in runs:
            assert all(i["status"] == D.OK for i in infos)
            assert np.array_equal(p, runs[0][0])
        sols.append(runs[0][0])
        s.close()
    assert np.array_equal(sols[0], sols[1])


def test_block_tuning_is_separate_and_pinning_covers_every_block(ctx):
    """dda_op_tune with 2 or more right-hand sides tunes a set of its own for
    blocks: the set for single right-hand sides is unchanged, a tuned block's
    every result matches the oracle, and pinned variants serve every block size
    again (a block of 8 then gives bitwise each right-hand side's result alone)."""
    g, params, einc, ccs = _setup("ellipsoid_5_4_3", 4, 12, (32, 64, 96))
    k, n = len(ccs), params.vec_size
    assert k == 8
    rng = np.random.default_rng(6)
    x = (rng.standard_normal((k, n)) + 1j * rng.standard_normal((k, n))).astype(np.complex64)
    for r in range(k):
        x[r, params.nv:params.vec_stride] = 0
    op = _op(ctx, g, params, 32)
    op.set_couple_constants_batch(ccs)
    print("tune 1:", op.tune(1), flush=True)
    v1 = op.variants()
    print("tune 8:", op.tune(8), flush=True)
    assert op.variants() == v1
    xb, yb = D.DeviceVector(ctx, k * n, 32), D.DeviceVector(ctx, k * n, 32)
    xb.upload(x.reshape(-1))
    op.apply_batch(xb, yb, k)
    yk = yb.download().reshape(k, n)
    for r in (0, k - 1):
        orc = cases.oracle_problem(g, params, ccs[r])
        assert cases.rel_err(yk[r].astype(np.complex128), orc.matvec(x[r].astype(np.complex128), precision=80)) <= 1e-5
    op.variants(v1)
    op.apply_batch(xb, yb, k)
    yk = yb.download().reshape(k, n)
    xs, ys = D.DeviceVector(ctx, n, 32), D.DeviceVector(ctx, n, 32)
    for r in range(k):
        op.set_couple_constants(ccs[r])
        xs.upload(x[r])
        op.apply(xs, ys)
        assert np.array_equal(yk[r], ys.download()), r
    op.close()


def test_mixed_block_solve_over_tuned_operator(ctx):
    """A mixed-mode block of 8 over a tuned operator (its FP32 twin then tuned
    for blocks at the first block solve) converges and matches the same block over
    an untuned operator to the solver's tolerance."""
    g, params, einc, ccs = _setup("ellipsoid_5_4_3", 4, 13, (32, 64, 96))
    k = len(ccs)
    pols = {}
    for tuned in (False, True):
        op = _op(ctx, g, params, 64)
        op.set_couple_constants(ccs[0])
        if tuned:
            op.tune(1)
        op.set_couple_constants_batch(ccs)
        s = D.GpuBicgStab(op, mixed=True)
        p, infos = s.get_pol_vecs(einc, 1e-8)
        assert all(i["status"] == D.OK for i in infos)
        pols[tuned] = p
        s.close()
        op.close()
    for r in range(k):
        assert cases.rel_err(pols[True][r], pols[False][r]) <= 1e-6
