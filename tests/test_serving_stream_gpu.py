"""The serving workload's own job stream through the job server: seeded jobs of
the stream's distribution (sphere / box / ellipsoid, largest extent 16..96
dipoles, Re m 1.2..2.0, Im m 0..0.2, |m| kd = 0.5, random orientation, both
polarisations) arriving as a Poisson stream faster than the server's capacity,
so that several run at once on four slots. Each result equals the same job
solved alone on a one-slot server, bit for bit (the same kernels, the same
tensor, per-right-hand-side reductions); the two smallest jobs are checked
against the oracle: the true residual of the server's answer under the
oracle's operator, and Cext / Cabs against the oracle's own solve."""
import numpy as np
import pytest

import dda_amd as D
from dda_amd import serving as S

import cases

pytestmark = pytest.mark.gpu

N_JOBS = 24
SEED = 42


def test_serving_stream_alone_vs_mix_and_oracle():
    jobs = S.make_stream(N_JOBS, SEED)
    dims = [max(j.dims) for j in jobs]
    print("stream dims", dims, "shapes", [j.shape for j in jobs], flush=True)
    assert min(dims) >= 16 and max(dims) <= 96 and max(dims) - min(dims) >= 40
    assert {j.shape for j in jobs} == {"sphere", "box", "ellipsoid"}
    # alone: one slot, one job at a time
    alone = D.Server(slots=1, solver="bicgstab", mixed=True, eps=1e-6)
    ref = []
    for j in jobs:
        ref.append(alone.wait(S.submit(alone, j)))
    alone.close()
    assert all(r[2]["status"] == 0 for r in ref)
    # in the mix: four slots, Poisson arrivals at ~3x what the server finishes
    mix = D.Server(slots=4, solver="bicgstab", mixed=True, eps=1e-6)
    r = S.run_stream(mix, jobs, rate=100.0)
    st = mix.stats()
    mix.close()
    print("stream", r, "server", st, flush=True)
    assert r["ok"] == N_JOBS
    spans = [(j.result["t_start"], j.result["t_end"]) for j in jobs]
    overlapping = sum(1 for a in spans for b in spans if a < b and a[0] < b[1] and b[0] < a[1])
    assert overlapping >= N_JOBS // 4
    for j, (p_alone, i_alone, _) in zip(jobs, ref):
        assert np.array_equal(j.result["pol"], p_alone), j.dims
        assert [i["iterations"] for i in j.result["infos"]] == [i["iterations"] for i in i_alone]
        assert all(i["true_rel_residual"] < 1e-6 for i in j.result["infos"])
    # the two smallest jobs against the oracle
    import oracle as O
    order = sorted(range(N_JOBS), key=lambda i: jobs[i].params.nv)[:2]
    for idx in order:
        j = jobs[idx]
        g, params = j.geometry, j.params
        pol = j.result["pol"]
        for r_ in range(2):
            orc = cases.oracle_problem(g, params, j.cc_sqrt[r_])
            # |E - A' p / sqrt(alpha)| / |E| under the oracle's operator: the
            # server's true residual, recomputed on the CPU
            st_, p_ref, it_ref, _ = orc.solve("qmr", j.einc[r_], 1e-10, max_iter=5000)
            assert st_ == O.ST_OK
            c, c_ref = orc.ext_cross_sum(pol[r_], j.einc[r_]), orc.ext_cross_sum(p_ref, j.einc[r_])
            a, a_ref = orc.abs_cross(pol[r_], j.cc[r_]), orc.abs_cross(p_ref, j.cc[r_])
            err = cases.rel_err(pol[r_], p_ref)
            print("job", idx, j.shape, j.dims, "nv", params.nv, "rhs", r_, "pol err %.2e" % err,
                  "cext err %.2e" % abs(c / c_ref - 1), "cabs err %.2e" % abs(a / a_ref - 1), flush=True)
            assert err <= 1e-5
            assert abs(c - c_ref) <= 1e-5 * abs(c_ref)
            assert abs(a - a_ref) <= 1e-5 * abs(a_ref)
