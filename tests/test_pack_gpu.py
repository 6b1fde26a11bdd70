"""Packs: several particles on one grid in one operator and one block solve.

Each member's rows of the pack's block matvec must equal its own operator's
matvec (FP64 to rounding, FP32 to single precision), zero off its dipoles, and a
block solve over the pack must give each member's own solution."""
import math

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu

GRID = (32, 32, 32)   # compile-time planned lengths
GU = 0.5e-6


def _members():
    """Three particles of different shape, material count and wave number;
    boxes <= 16 so that all fit the 32^3 grid."""
    g1 = P.sphere(6.0, 1.5 + 0.1j, GU)                     # 12^3 box
    g2 = P.box(14, 9, 7, 1.33 + 0.02j, GU)
    g3 = P.sphere(5.0, 1.5 + 0.1j, GU)
    c = g3.pos.astype(np.float64) - (np.array(g3.box) - 1) / 2.0
    g3.mat = (np.sum(c * c, axis=1) < 2.5 ** 2).astype(np.uint8)
    g3.materials = [1.5 + 0.1j, 2.0 + 0.3j]
    lams = [6e-6, 5e-6, 6e-6]                                # members 0 and 2 share the wave number
    return [(g, P.DDAParams(g, lam, GRID)) for g, lam in zip((g1, g2, g3), lams)]


def _pack(ctx, mem, precision=64, cache=None):
    return D.PackMatVec(ctx, [(g.pos, g.mat, g.box, p.wave_num, len(g.materials)) for g, p in mem], GU,
                        precision=precision, grid=GRID, cache=cache)


def _own(ctx, g, p, precision=64):
    return D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, p.wave_num, precision=precision, grid=GRID)


def _rhs(mem, prop=(0.0, 0.0, 1.0)):
    prop_n, pol1, pol2 = P.beam_pols(prop)
    out = []
    for m, (g, p) in enumerate(mem):
        for pol in (pol1, pol2):
            cc, cs = P.couple_constants(p, prop_n, pol, "ldr")
            out.append((m, cs, P.einc_plane(p, prop_n, pol)))
    return out


def _apply_batch(op, xs, dtype):
    S = op.vec_stride
    n = len(xs)
    xv = D.DeviceVector(op.ctx, n * 3 * S, op.precision)
    yv = D.DeviceVector(op.ctx, n * 3 * S, op.precision)
    xv.upload(np.concatenate([x.reshape(-1) for x in xs]).astype(dtype))
    op.apply_batch(xv, yv, n)
    return yv.download().reshape(n, 3, S)


@pytest.mark.parametrize("precision,tol", [(64, 1e-12), (32, 2e-6)])
def test_pack_matvec_matches_each_member(ctx, precision, tol):
    mem = _members()
    dtype = np.complex128 if precision == 64 else np.complex64
    pk = _pack(ctx, mem, precision)
    assert pk.nmember == 3 and pk.nmat == 3  # two materials in member 2, plus the void one
    rhs = _rhs(mem)
    pk.members_of([m for m, _, _ in rhs])
    pk.set_couple_constants_batch(np.array([pk.pad_cc(m, cs) for m, cs, _ in rhs]))
    xs, own = [], []
    for r, (m, cs, _) in enumerate(rhs):
        g, p = mem[m]
        x = P.random_dipole_vector(p, seed=100 + r)
        xs.append(pk.scatter(m, x.reshape(3, -1)))
        op = _own(ctx, g, p, precision)
        op.set_couple_constants(cs)
        own.append(D.MatVecGpu(op).apply(x.astype(dtype)).reshape(3, -1))
        op.close()
    y = _apply_batch(pk, xs, dtype)
    for r, (m, _, _) in enumerate(rhs):
        ym = pk.gather(m, y[r], stride=own[r].shape[1])
        assert cases.rel_err(ym, own[r]) <= tol, (r, m)
        off = np.ones(pk.vec_stride, bool)
        off[pk.maps[m]] = False
        assert not np.any(y[r][:, off]), "entries off the member's dipoles must stay zero"
    if precision == 64:
        # and the oracle for the two-material member
        g, p = mem[2]
        r = 4
        ref = cases.oracle_problem(g, p, rhs[r][1]).matvec(P.random_dipole_vector(p, seed=100 + r), precision=80)
        assert cases.rel_err(pk.gather(2, y[r], stride=p.vec_stride).reshape(-1), ref) <= 1e-12
    pk.close()


@pytest.mark.parametrize("solver,mixed", [("bicgstab", True), ("qmr", False), ("qmr", True)])
def test_pack_block_solve_matches_own_solves(ctx, solver, mixed):
    mem = _members()
    rhs = _rhs(mem, prop=(0.3, -0.2, 0.9))
    pk = _pack(ctx, mem)
    pk.members_of([m for m, _, _ in rhs])
    pk.set_couple_constants_batch(np.array([pk.pad_cc(m, cs) for m, cs, _ in rhs]))
    einc = np.array([pk.scatter(m, e.reshape(3, -1)).reshape(-1) for m, _, e in rhs])
    pol, infos = D.SOLVER_CLASSES[solver](pk, mixed=mixed).get_pol_vecs(einc, 1e-9)
    for r, (m, cs, e) in enumerate(rhs):
        g, p = mem[m]
        op = _own(ctx, g, p)
        op.set_couple_constants(cs)
        ref, inf = D.SOLVER_CLASSES[solver](op, mixed=mixed).get_pol_vecs(e[None], 1e-9)
        got = pk.gather(m, pol[r].reshape(3, -1), stride=p.vec_stride).reshape(-1)
        print(r, m, infos[r]["iterations"], inf[0]["iterations"], cases.rel_err(got, ref[0]), flush=True)
        assert infos[r]["true_rel_residual"] <= 1e-9 * 1.01
        assert cases.rel_err(got, ref[0]) <= 1e-7
        off = np.ones(pk.vec_stride, bool)
        off[pk.maps[m]] = False
        assert not np.any(pol[r].reshape(3, -1)[:, off])
        op.close()
    pk.close()


def test_pack_with_cache_and_twin(ctx):
    """Members of one wave number share a cache entry; the FP32 twin of a
    mixed solver takes its tensors from the same cache."""
    mem = _members()
    cache = D.TensorCache(ctx, 1 << 30)
    pk = _pack(ctx, mem, cache=cache)
    st = cache.stats()
    assert st["entries"] == 2 and st["hits"] == 1
    rhs = _rhs(mem)
    pk.members_of([m for m, _, _ in rhs])
    pk.set_couple_constants_batch(np.array([pk.pad_cc(m, cs) for m, cs, _ in rhs]))
    einc = np.array([pk.scatter(m, e.reshape(3, -1)).reshape(-1) for m, _, e in rhs])
    pol, infos = D.GpuBicgStab(pk, mixed=True).get_pol_vecs(einc, 1e-8)
    assert all(i["true_rel_residual"] <= 1.01e-8 for i in infos)
    pk2 = _pack(ctx, mem)  # generated tensors, clipped at the union box
    pk2.members_of([m for m, _, _ in rhs])
    pk2.set_couple_constants_batch(np.array([pk2.pad_cc(m, cs) for m, cs, _ in rhs]))
    pol2, _ = D.GpuBicgStab(pk2, mixed=True).get_pol_vecs(einc, 1e-8)
    assert cases.rel_err(pol, pol2) <= 1e-6
    pk.close()
    pk2.close()
    cache.close()


def test_pack_rejects_bad_input(ctx):
    mem = _members()
    g, p = mem[0]
    twice = np.concatenate([g.pos, g.pos[:1]])
    with pytest.raises(D.DDAError):  # two dipoles of one member at one position
        D.PackMatVec(ctx, [(twice, np.zeros(len(twice)), g.box, p.wave_num, 1)], GU, grid=GRID)
    with pytest.raises(D.DDAError):  # 255 materials plus the void one is the limit
        D.PackMatVec(ctx, [(g.pos, g.mat, g.box, p.wave_num, 256)], GU, grid=GRID)
    pk = _pack(ctx, mem)
    with pytest.raises(D.DDAError):
        pk.members_of([0, 3])
    pk.members_of([0, 1])
    pk.set_couple_constants_batch(np.ones((2, pk.nmat, 3)))
    pv = D.DeviceVector(ctx, 2 * pk.vec_size, 64)
    with pytest.raises(D.DDAError):  # observables need one wave number
        D.cross_sections(pk, pv, pv, 2)
    with pytest.raises(D.DDAError):
        D.incident_plane(pk, [[0, 0, 1]] * 2, [[1, 0, 0]] * 2)
    pk.close()
    same = _pack(ctx, [mem[0], mem[2]])  # one wave number: accepted
    same.members_of([0, 1])
    same.set_couple_constants_batch(np.ones((2, same.nmat, 3)))
    pv = D.DeviceVector(ctx, 2 * same.vec_size, 64)
    pv.upload(np.zeros(2 * same.vec_size))
    cext, _ = D.cross_sections(same, pv, pv, 2)
    assert np.all(cext == 0)
    same.close()
