"""Tensor cache and job server: results independent of the mix, admission by
footprint, cache reuse across geometries on one grid and kd."""
import math
import threading
import time

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P
from dda_amd import serving as S

import cases

pytestmark = pytest.mark.gpu


def test_cached_tensor_matches_box_clipped(ctx):
    """The cache clips the Green's function at G/2 instead of the box: the
    product is the same inside the box (FP64, to rounding), and a second
    geometry on the same grid and kd reuses the entry."""
    g, params, cc, cs, prop, pol = cases.build("ellipsoid_5_4_3", grid=(24, 20, 16))
    cache = D.TensorCache(ctx, 1 << 30)
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid)
    opc = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, grid=params.grid, cache=cache)
    op.set_couple_constants(cs)
    opc.set_couple_constants(cs)
    x = P.random_dipole_vector(params, seed=3)
    y, yc = D.MatVecGpu(op).apply(x), D.MatVecGpu(opc).apply(x)
    assert cases.rel_err(yc, y) <= 1e-13
    # the oracle agrees too
    assert cases.rel_err(yc, cases.oracle_problem(g, params, cs).matvec(x, precision=80)) <= 1e-12
    g2 = P.box(6, 5, 4, 1.33 + 0.01j, g.grid_unit)
    op2 = D.GpuMatVec(ctx, g2.pos, g2.mat, g2.box, g2.grid_unit, params.wave_num, grid=params.grid, cache=cache)
    st = cache.stats()
    assert st["hits"] == 1 and st["misses"] == 1 and st["entries"] == 1
    op2.set_couple_constants(cs)
    x2 = P.random_dipole_vector(P.DDAParams(g2, 2 * math.pi / params.wave_num, params.grid), seed=4)
    ref2 = cases.oracle_problem(g2, P.DDAParams(g2, 2 * math.pi / params.wave_num, params.grid), cs)
    assert cases.rel_err(D.MatVecGpu(op2).apply(x2), ref2.matvec(x2, precision=80)) <= 1e-12
    for o in (op, opc, op2):
        o.close()
    cache.close()


def _small_jobs():
    rng = np.random.default_rng(7)
    j1 = S.make_job(rng, "sphere", 12, 1.5 + 0.1j)
    j2 = S.make_job(rng, "box", 14, 1.33 + 0.05j)
    # the particle of j1 in another orientation: packed with j1 when both wait
    j3 = S.make_job(np.random.default_rng(8), "sphere", 12, 1.5 + 0.1j)
    return j1, j2, j3


@pytest.mark.parametrize("solver,mixed", [("bicgstab", True), ("qmr", False)])
def test_server_result_same_alone_and_in_mix(solver, mixed):
    j1, j2, j3 = _small_jobs()
    with_alone = D.Server(slots=1, solver=solver, mixed=mixed, eps=1e-8)
    pol_alone, infos_alone, res = with_alone.wait(S.submit(with_alone, j1))
    assert res["status"] == 0 and res["packed_rhs"] == 2
    with_alone.close()
    mix = D.Server(slots=3, solver=solver, mixed=mixed, eps=1e-8)
    ids = [S.submit(mix, j) for j in (j2, j1, j3)]
    out = [mix.wait(i) for i in ids]
    st = mix.stats()
    mix.close()
    assert all(o[2]["status"] == 0 for o in out)
    pol_mix, infos_mix, res_mix = out[1]
    print("packed", [o[2]["packed_rhs"] for o in out], "stats", st, flush=True)
    # bitwise: same kernels, same cached tensor, per-RHS reductions
    assert np.array_equal(pol_mix, pol_alone)
    assert [i["iterations"] for i in infos_mix] == [i["iterations"] for i in infos_alone]
    # and the same as a direct solve through the plain API (box-clipped tensor)
    ctx = D.Context(0)
    g = j1.geometry
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, j1.params.wave_num, grid=j1.params.grid)
    op.set_couple_constants_batch(j1.cc_sqrt)
    p_direct, _ = D.SOLVER_CLASSES[solver](op, mixed=mixed).get_pol_vecs(j1.einc, 1e-8)
    assert cases.rel_err(pol_alone, p_direct) <= 1e-6
    ctx.close()


def test_server_second_waiter_is_refused():
    """One thread waits per job: of two concurrent waits on one job, one
    collects it and the other returns INVALID_ARG (never a freed job)."""
    j1, _, _ = _small_jobs()
    srv = D.Server(slots=1, eps=1e-6)
    jid = S.submit(srv, j1)
    L = D.lib()
    first = []
    t = threading.Thread(target=lambda: first.append(L.dda_server_wait(srv.handle, jid, None, None, None)))
    t.start()
    time.sleep(0.05)
    second = L.dda_server_wait(srv.handle, jid, None, None, None)
    t.join()
    srv._shapes.pop(jid)
    srv.close()
    assert sorted([first[0], second]) == sorted([D.OK, D.INVALID_ARG])


def test_server_admission_and_cache():
    """A budget that holds one job at a time serialises the jobs; all finish;
    jobs on the same grid and kd share the cached tensor."""
    rng = np.random.default_rng(11)
    # box 10x8x6 and the ellipsoid in it: different particles, one grid and kd
    jobs = [S.make_job(rng, "box", 10, 1.4 + 0.02j, other=(8, 6)) for _ in range(2)]
    jobs += [S.make_job(rng, "ellipsoid", 10, 1.4 + 0.02j, other=(8, 6)) for _ in range(2)]
    assert jobs[0].params.grid == jobs[2].params.grid
    one = D.Server(slots=4, max_pack=2, eps=1e-6)
    foot = one.wait(S.submit(one, jobs[0]))[2]["footprint"]
    one.close()
    srv = D.Server(slots=4, max_pack=2, eps=1e-6, hbm_budget=int(foot * 1.5) + (1 << 26), cache_bytes=1 << 26)
    ids = [S.submit(srv, j) for j in jobs]
    out = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    print(st, [(o[2]["t_start"], o[2]["t_end"]) for o in out], flush=True)
    assert all(o[2]["status"] == 0 for o in out)
    assert st["peak_bytes"] <= foot * 1.5
    spans = sorted((o[2]["t_start"], o[2]["t_end"]) for o in out)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))  # one at a time
    assert st["cache_hits"] >= 1  # box and ellipsoid of the same grid and kd share the tensor


def test_server_stream_overlaps():
    """Continuous batching on a short Poisson stream: every job completes,
    several run at once, and the latencies are reported."""
    rng = np.random.default_rng(5)
    jobs = [S.make_job(rng, dim=int(rng.integers(12, 40))) for _ in range(10)]
    srv = D.Server(slots=4, eps=1e-6)
    r = S.run_stream(srv, jobs, rate=200.0)
    srv.close()
    print(r, flush=True)
    assert r["ok"] == len(jobs)
    spans = [(j.result["t_start"], j.result["t_end"]) for j in jobs]
    overlapping = sum(1 for a in spans for b in spans if a < b and a[0] < b[1] and b[0] < a[1])
    assert overlapping >= 1
    assert r["p99_ms"] >= r["p50_ms"] > 0


@pytest.mark.parametrize("solver,mixed", [("bicgstab", True), ("qmr", False)])
def test_server_packs_particles_on_one_grid(solver, mixed):
    """pack_grids: queued jobs of different particles on one grid run as one
    block solve over a pack operator; each result equals the job's solve alone
    to the solver's tolerance."""
    rng = np.random.default_rng(21)
    jobs = [S.make_job(rng, "box", 10, 1.4 + 0.02j, other=(8, 6)),
            S.make_job(rng, "ellipsoid", 10, 1.6 + 0.05j, other=(8, 6)),
            S.make_job(rng, "sphere", 12, 1.5 + 0.1j),
            S.make_job(rng, "box", 9, 1.3 + 0.01j, other=(7, 5))]
    alone = D.Server(slots=1, solver=solver, mixed=mixed, eps=1e-9)
    ref = [alone.wait(S.submit(alone, j)) for j in jobs]
    alone.close()
    assert all(r[2]["status"] == 0 and r[2]["packed_rhs"] == 2 for r in ref)
    srv = D.Server(slots=1, solver=solver, mixed=mixed, eps=1e-9, pack_grids=True)
    ids = [S.submit(srv, j) for j in jobs]
    out = [srv.wait(i) for i in ids]
    srv.close()
    print([o[2]["packed_rhs"] for o in out], flush=True)
    assert all(o[2]["status"] == 0 for o in out)
    assert max(o[2]["packed_rhs"] for o in out) >= 6  # at least three particles in one solve
    for (p_ref, i_ref, _), (p, infos, _) in zip(ref, out):
        assert cases.rel_err(p, p_ref) <= 1e-7
        assert all(i["true_rel_residual"] <= 1.01e-9 for i in infos)
        # padding entries stay zero
        assert p.shape == p_ref.shape
