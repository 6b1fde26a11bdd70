#!/usr/bin/env python3
"""Benchmark: DDA solves/s at relative residual 1e-6 on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): a sphere 128
dipoles across (1,099,136 dipoles), m = 1.5 + 0.1i, |m| kd = 0.5, FFT grid
256^3, plane wave along z, both polarisations solved to |r| / |b| < 1e-6. One
step = the two solves. Runs: FP64 QMR-CS (the baseline), FP64 BiCGStab, mixed
QMR-CS, mixed BiCGStab, and the two mixed ones with both polarisations as one
2-RHS block; value = the last run. Inputs are synthetic (geometry and incident
fields generated here) and resident on the GPU when the timed region starts;
the polarisations stay there.

Printed JSON line: value = solves/s; `roofline` for the dominant matvec kernel
from HIP events recorded around every launch in the timed region; the
whole-matvec rate against the algorithmic bytes
B_mv = c [9 nv + 12 Gx By Bz + 12 Gx By Gz + 6 Gx Gy Gz]; `cpu_baseline` = the
reference's CPU matvec (MatVecCpu) restated over scipy.fft, timed on this
host's threads on a bounded sample (a few matvecs) and scaled to one solve.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "DDA solves/s at resid 1e-6 (mixed vs FP64, same GPU); matvec HBM GB/s; p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
PASS_NAMES = ["k_pass1 (X-FFT + scatter)", "k_colfft (Y-FFT)", "k_pass3 (Z-FFT, D, inverse Z)",
              "k_colfft (inverse Y)", "k_pass5 (inverse X + gather)"]


def pass_bytes(nv, S, box, grid, n_rows_nonempty, c, nrhs=1):
    """Algorithmic HBM bytes of each of the five kernels for one matvec over
    nrhs right-hand sides (the octant tensor is read once per launch)."""
    bx, by, bz = box
    gx, gy, gz = grid
    nrows = by * bz
    idx = 8 * nv + 4 * (nrows + 1)
    oct_d = 6 * (gx // 2 + 1) * (gy // 2 + 1) * (gz // 2 + 1) * c
    a1 = 3 * bz * by * gx * c
    a2 = 3 * bz * gy * gx * c
    return [
        nrhs * (3 * nv * c + a1) + idx + nv,             # x, index, materials -> A1
        nrhs * (a1 + a2),                                # A1 -> A2
        nrhs * 2 * a2 + oct_d,                           # A2 (+ octant D) -> A2
        nrhs * (a2 + a1),                                # A2 -> A1
        nrhs * (3 * n_rows_nonempty * gx * c + 3 * nv * c + 3 * S * c) + idx + nv,  # A1, x -> y
    ]


def b_mv(nv, box, grid, c, nrhs=1):
    """B_mv of one matvec; for a block of nrhs the tensor term counts once."""
    bx, by, bz = box
    gx, gy, gz = grid
    return c * (nrhs * (9 * nv + 12 * gx * by * bz + 12 * gx * by * gz) + 6 * gx * gy * gz)


def measured_traffic(diameter, grid, prec, pass_index, nrhs=1):
    """HBM bytes per launch of one matvec kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy of known size),
    or None when no measurement exists for this workload."""
    path = os.path.join(ROOT, "profiles", "traffic_matvec_sphere128.md")
    if diameter != 128 or tuple(grid) != (256, 256, 256) or not os.path.exists(path):
        return None
    text = open(path).read()
    blob = text.split("```json", 1)[1].split("```", 1)[0]
    data = json.loads(blob)
    key = prec if nrhs == 1 else "%s_rhs%d" % (prec, nrhs)
    if key not in data:
        return None
    return data[key][pass_index]["traffic_bytes"]


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def matvec_profile(ctx, op64, g, params, rhs, k, D, P, np, nrhs=1):
    """Copy 512 MiB device to device once, then k matvecs in FP32 and in FP64
    (each over nrhs right-hand sides as one block launch per pass)."""
    n = (512 << 20) // 16
    a, b = D.DeviceVector(ctx, n, 64), D.DeviceVector(ctx, n, 64)
    b.copy_from(a)
    out = {"copy_bytes": 512 << 20, "nrhs": nrhs}
    op32 = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=32, grid=params.grid)
    for name, op in (("fp32", op32), ("fp64", op64)):
        op.set_couple_constants_batch(np.array([rhs[r % len(rhs)][2] for r in range(nrhs)]))
        out[name + "_tuning"] = op.tune(nrhs)  # the kernel variants the solves run
        prec = 32 if name == "fp32" else 64
        x = D.DeviceVector(ctx, nrhs * params.vec_size, prec)
        y = D.DeviceVector(ctx, nrhs * params.vec_size, prec)
        x.upload(np.concatenate([P.random_dipole_vector(params, seed=r + 1) for r in range(nrhs)]))
        op.apply_batch(x, y, nrhs)
        ctx.synchronize()
        op.profile(True)
        for _ in range(k):
            op.apply_batch(x, y, nrhs)
        ms, cnt, _ = op.profile_read()
        op.profile(False)
        out[name + "_pass_avg_us"] = [1e3 * ms[i] / max(cnt[i], 1) for i in range(5)]
    print(json.dumps(out))
    return 0


def matvec_leg(ctx, op64, g, params, rhs, D, P, np):
    """Per-apply times of the workload's matvec: FP64 and FP32 with one
    right-hand side, FP32 with both polarisations as one block. x: nv entries
    per component with re, im ~ U[-1, 1) (numpy default_rng, seed 0x5EED)."""
    out = {"method": "3 warm-up applies, then 20 applies, each timed as the sum of its five kernels' "
                     "HIP-event durations on the operator's stream"}
    op32 = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=32, grid=params.grid)
    try:
        op32.set_couple_constants_batch(np.array([r[2] for r in rhs]))
        out["fp32_tuning"] = op32.tune(2)
        x1 = P.random_dipole_vector(params)
        x2 = np.concatenate([x1, P.random_dipole_vector(params, seed=0x5EED + 1)])
        for name, op, prec, nrhs in (("fp64_1rhs", op64, 64, 1), ("fp32_1rhs", op32, 32, 1),
                                     ("fp32_2rhs", op32, 32, 2)):
            x = D.DeviceVector(ctx, nrhs * op.vec_size, prec)
            y = D.DeviceVector(ctx, nrhs * op.vec_size, prec)
            x.upload((x2 if nrhs == 2 else x1).astype(np.complex128 if prec == 64 else np.complex64))
            if op is op64:
                op.set_couple_constants(rhs[0][2])
            apply = (lambda: op.apply_batch(x, y, nrhs)) if nrhs > 1 else (lambda: op.apply(x, y))
            for _ in range(3):
                apply()
            ctx.synchronize()
            op.profile(True)
            op.profile_read()
            us = []
            for _ in range(20):
                apply()
                ctx.synchronize()
                ms, cnt, _ = op.profile_read()
                us.append(1e3 * float(np.sum(ms)))
            op.profile(False)
            x.close()
            y.close()
            out[name] = {"median_us": float(np.median(us)), "p10_us": float(np.percentile(us, 10)),
                         "p90_us": float(np.percentile(us, 90))}
    finally:
        op32.close()
    return out


def c3_bench(args, D, P, np, ctx):
    """Batched right-hand sides: ellipsoid with semi-axes 32x24x16 dipoles,
    m = 1.33+0.01i, |m| kd = 0.6, 64 seeded directions x 2 polarisations = 128
    fields sharing one tensor. Mixed block QMR-CS and mixed block BiCGStab (one
    128-RHS block each) against FP64 QMR-CS solving the same fields one at a
    time, all to args.eps on device-resident fields; Cext of every field
    compared."""
    m, d = 1.33 + 0.01j, 1e-7
    g = P.ellipsoid(32, 24, 16, m, d)
    params = P.DDAParams(g, 2 * math.pi * d * abs(m) / 0.6)
    einc, ccs, props, pols = P.incident_batch(params, P.directions(64, 2024))
    k = len(ccs)
    op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, g.grid_unit, params.wave_num, precision=64, grid=params.grid)
    op.set_couple_constants(ccs[0])
    if not args.no_tune:
        op.tune(1)
    vs = op.vec_size
    e1, p1 = D.DeviceVector(ctx, vs, 64), D.DeviceVector(ctx, vs, 64)
    s64 = D.GpuQmrCs(op)
    e1.upload(einc[0])
    s64.solve_device(e1, p1, 1, args.eps)  # warm-up
    ctx.synchronize()
    p64, it64, el64 = [], [], 0.0
    for r in range(k):
        e1.upload(einc[r])  # outside the timed span
        ctx.synchronize()
        t0 = time.perf_counter()
        s64.set_couple_constants(ccs[r])
        inf = s64.solve_device(e1, p1, 1, args.eps)[0]
        ctx.synchronize()
        el64 += time.perf_counter() - t0
        p64.append(p1.download())
        it64.append(int(inf["iterations"]))
    s64.close()
    e1.close()
    p1.close()
    ek, pk = D.DeviceVector(ctx, k * vs, 64), D.DeviceVector(ctx, k * vs, 64)
    ek.upload(einc.reshape(-1))
    out = {"workload": "ellipsoid semi-axes 32x24x16 dipoles (nv=%d), m=1.33+0.01i, |m|kd=0.6, grid %s, 64 seeded "
                       "directions x 2 polarisations = %d right-hand sides, to %g" % (
                           params.nv, "x".join(map(str, params.grid)), k, args.eps),
           "fp64_qmr_one_at_a_time": {"solves_per_s": k / el64, "seconds": el64,
                                      "iterations": [min(it64), max(it64)]}}
    for name, cls in (("mixed_qmr_block", D.GpuQmrCs), ("mixed_bicgstab_block", D.GpuBicgStab)):
        sb = cls(op, mixed=True)
        op.set_couple_constants_batch(ccs)
        sb.solve_device(ek, pk, k, args.eps)  # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        op.set_couple_constants_batch(ccs)
        infos = sb.solve_device(ek, pk, k, args.eps)
        ctx.synchronize()
        elb = time.perf_counter() - t0
        pb = pk.download().reshape(k, vs)
        sb.close()
        cerr = max(abs(P.ext_cross(params, pb[r], einc[r]) / P.ext_cross(params, p64[r], einc[r]) - 1)
                   for r in range(k))
        out[name] = {"solves_per_s": k / elb, "seconds": elb, "vs_fp64_qmr": el64 / elb,
                     "iterations": [min(int(i["iterations"]) for i in infos), max(int(i["iterations"]) for i in infos)],
                     "ok": sum(1 for i in infos if i["status"] == D.OK), "cext_max_rel_err": cerr}
    ek.close()
    pk.close()
    op.close()
    return out


def large_matvec(args, D, P, np, ctx, solvers=()):
    """The largest single-GPU grid: sphere 512 dipoles across (~70M dipoles,
    1024^3), m = 1.6+0.05i, |m| kd = 0.5. FP32 and FP64 matvec rates against
    B_mv, the device memory held, and mixed solves to 1e-5 with the given
    solvers ("bicgstab", "qmr"; one polarisation each), their Cext and Cabs
    beside Mie theory for the volume-equivalent sphere."""
    m = 1.6 + 0.05j
    d = 1e-7
    kd = 0.5 / abs(m)
    t0 = time.perf_counter()
    g = P.sphere(256.0, m, d)
    params = P.DDAParams(g, 2 * math.pi * d / kd)
    t_geom = time.perf_counter() - t0
    print("large: nv=%d grid=%s geometry %.1f s" % (params.nv, params.grid, t_geom), file=sys.stderr, flush=True)
    prop, pol1, pol2 = P.beam_pols((0, 0, 1))
    cc, cs = P.couple_constants(params, prop, pol1)
    out = {"workload": "sphere 512 dipoles across (nv=%d), m=1.6+0.05i, |m|kd=0.5, grid %s"
                       % (params.nv, "x".join(map(str, params.grid)))}
    rows = np.unique(g.pos[:, 2].astype(np.int64) * g.box[1] + g.pos[:, 1]).size
    xh = P.random_dipole_vector(params)
    op64 = None
    for name, prec in (("fp32", 32), ("fp64", 64)):
        t0 = time.perf_counter()
        op = D.GpuMatVec(ctx, g.pos, g.mat, g.box, d, params.wave_num, precision=prec, grid=params.grid)
        op.set_couple_constants(cs)
        t_create = time.perf_counter() - t0
        tuning = op.tune(1) if not args.no_tune else "not tuned"
        x = D.DeviceVector(ctx, params.vec_size, prec)
        y = D.DeviceVector(ctx, params.vec_size, prec)
        x.upload(xh)
        op.apply(x, y)
        ctx.synchronize()
        # the rate as a caller sees it (no events), then per-pass event times
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            op.apply(x, y)
        ctx.synchronize()
        wall = (time.perf_counter() - t0) / reps * 1e3
        op.profile(True)
        for _ in range(reps):
            op.apply(x, y)
        ms, cnt, _ = op.profile_read()
        op.profile(False)
        pav = [ms[i] / max(cnt[i], 1) for i in range(5)]
        mv = float(sum(pav))
        c = 8 if prec == 32 else 16
        pb = pass_bytes(params.nv, params.vec_stride, g.box, params.grid, rows, c)
        bmv = b_mv(params.nv, g.box, params.grid, c)
        out[name] = {"matvec_ms": wall, "matvec_ms_events": mv, "pass_avg_us": [1e3 * v for v in pav],
                     "pass_gbs": [pb[i] / (pav[i] * 1e-3) / 1e9 for i in range(5)],
                     "gbs_vs_bmv": bmv / (wall * 1e-3) / 1e9, "frac_vs_bmv": bmv / (wall * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "gbs_algorithmic": sum(pb) / (wall * 1e-3) / 1e9, "device_bytes": op.device_bytes,
                     "create_s": t_create, "tuning": tuning, "variants": op.variants()}
        print("large %s: %s" % (name, json.dumps(out[name])), file=sys.stderr, flush=True)
        x.close()
        y.close()
        if prec == 64:
            op64 = op
        else:
            op.close()
    if solvers:
        einc = P.einc_plane(params, prop, pol1)
        out["solves"] = {}
        mie = None
        try:
            import oracle as O  # the checker of the cross sections only
            a_eq = d * (3.0 * params.nv / (4.0 * math.pi)) ** (1.0 / 3.0)
            qext, qsca, _, _ = O.mie(params.wave_num * a_eq, m)
            area = math.pi * a_eq * a_eq
            mie = (qext * area, (qext - qsca) * area)
        except Exception as e:
            out["mie_error"] = str(e)
        for kind in solvers:
            s = D.SOLVER_CLASSES[kind](op64, mixed=True, max_iter=args.large_max_iter)
            t0 = time.perf_counter()
            try:
                p = s.get_pol_vec(einc, 1e-5)
                status = "ok"
            except D.DDAError as e:
                p, status = None, str(e)
            el = time.perf_counter() - t0
            info = dict(s.info)
            res = {"seconds": el, "status": status, "iterations": info["iterations"],
                   "refinements": info["refinements"], "true_rel_residual": info["true_rel_residual"]}
            if p is not None:
                res["cext"] = P.ext_cross(params, p, einc)
                res["cabs"] = P.abs_cross(params, p, cc)
                if mie is not None:
                    res["cext_vs_mie"] = res["cext"] / mie[0] - 1
                    res["cabs_vs_mie"] = res["cabs"] / mie[1] - 1
            out["solves"][kind] = res
            print("large solve %s: %s" % (kind, json.dumps(res)), file=sys.stderr, flush=True)
            s.close()
    out["device_bytes_total"] = op64.device_bytes + out["fp32"]["device_bytes"]
    op64.close()
    return out


def large_bench(args, D, P, np):
    ctx = D.Context(0)
    kinds = args.large_solvers.split(",") if args.large_solves else ()
    print(json.dumps(large_matvec(args, D, P, np, ctx, solvers=kinds)))
    return 0


def _plain(o, path="line"):
    """The output line with numpy scalars and arrays as plain JSON values
    (complex ones as [re, im], reported on stderr: none is expected)."""
    if isinstance(o, dict):
        return {k: _plain(v, "%s.%s" % (path, k)) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_plain(v, "%s[%d]" % (path, i)) for i, v in enumerate(o)]
    if isinstance(o, np.ndarray):
        return _plain(o.tolist(), path)
    if isinstance(o, (complex, np.complexfloating)):
        print("bench: complex value at %s" % path, file=sys.stderr)
        return [float(o.real), float(o.imag)]
    if isinstance(o, np.generic):
        return o.item()
    return o


def serving_bench(args, D):
    """The serving stream: jobs of random shape (sphere / box / ellipsoid, largest
    extent 16..96 dipoles), Re m in 1.2..2, Im m in 0..0.2, |m| kd = 0.5, random
    orientation, both polarisations, mixed BiCGStab to eps, args.serve_slots
    slots. Capacity from a burst of jobs submitted at once; then fresh seeded
    streams with Poisson arrivals at 0.5x, 0.9x and 1.2x that capacity; latency
    = completion - submission. The streams are generated on host threads ahead
    of use. Runs in its own process (--serve-only), as a server would."""
    from concurrent.futures import ThreadPoolExecutor
    from dda_amd import serving as S
    loads = tuple(float(x) for x in args.serve_loads.split(","))
    out = {"workload": "seeded streams of %d jobs per load (burst of %d for capacity), sphere/box/ellipsoid, extent "
                       "U{16..96}, Re m U(1.2,2), Im m U(0,0.2), |m|kd=0.5, random orientation, 2 polarisations, "
                       "mixed BiCGStab to %g, %d slots, long solves' GPU share %s" % (
                           args.serve_jobs, args.serve_burst, args.eps, args.serve_slots, args.serve_share or "off")}
    # every stream is generated before any runs (no host threads competing
    # with the submitting one)
    with ThreadPoolExecutor(4) as ex:
        f_burst = ex.submit(S.make_stream, args.serve_burst, 42)
        f_streams = [ex.submit(S.make_stream, args.serve_jobs, 43 + i, None, "stream %d" % (43 + i))
                     for i in range(len(loads))]
        burst_jobs = f_burst.result()
        streams = [f.result() for f in f_streams]
    share = tuple(float(x) for x in args.serve_share.split(",")) if args.serve_share else None
    srv = D.Server(slots=args.serve_slots, solver="bicgstab", mixed=True, eps=args.eps, long_solve=share)
    try:
        S.run_stream(srv, S.make_stream(2, seed=7))  # warm-up: library and kernel load
        burst = S.run_stream(srv, burst_jobs, keep_pol=False, label="burst")
        cap = burst["jobs"] / burst["makespan_s"]
        out["capacity"] = dict(burst, capacity_jobs_per_s=cap)
        for i, load in enumerate(loads):
            r = S.run_stream(srv, streams[i], rate=load * cap, keep_pol=False, label="%.1fx" % load)
            out["poisson_%.1fx" % load] = dict(r, rate_jobs_per_s=load * cap)
            print("serving %.1fx: %s" % (load, json.dumps(r)), file=sys.stderr, flush=True)
            streams[i] = None
        out["server"] = srv.stats()
    finally:
        srv.close()
    # the burst again with different particles of one grid packed into one solve
    srv = D.Server(slots=args.serve_slots, solver="bicgstab", mixed=True, eps=args.eps, pack_grids=True)
    try:
        S.run_stream(srv, S.make_stream(2, seed=7))
        out["capacity_pack_grids"] = S.run_stream(srv, burst_jobs, keep_pol=False, label="burst, packed")
    finally:
        srv.close()
    ref = out.get("poisson_0.9x") or out["poisson_%.1fx" % loads[-1]]
    out.update({"p50_ms": ref["p50_ms"], "p99_ms": ref["p99_ms"], "solves_per_s": ref["solves_per_s"]})
    return out


def serving_child(args):
    """serving_bench in a child process of its own (a fresh HIP runtime with no
    other contexts or streams, as a serving process has)."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--serve-only", "--eps", repr(args.eps),
           "--serve-jobs", str(args.serve_jobs), "--serve-burst", str(args.serve_burst),
           "--serve-slots", str(args.serve_slots), "--serve-loads", args.serve_loads, "--serve-share",
           args.serve_share]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=600)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "serving child exited with %d" % r.returncode}
    return json.loads(lines[-1])


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--runs", default="fp64:qmr,fp64:bicgstab,mixed:qmr,mixed:bicgstab,mixed:qmr:batch,"
                                      "mixed:bicgstab:batch",
                    help="mode:solver[:batch] runs to time (batch = both polarisations in one block solve); "
                         "value = the last one, the FP64 QMR-CS run is the baseline")
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--diameter", type=int, default=128, help="sphere diameter in dipoles")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the 128-RHS ellipsoid batch figure")
    ap.add_argument("--no-matvec-applies", action="store_true",
                    help="skip the per-apply matvec timing (median, p10, p90)")
    ap.add_argument("--host-step", action="store_true",
                    help="also time one step through the host-array path (PCIe copies included)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default kernel variants)")
    ap.add_argument("--profile-rhs", type=int, default=1, help="with --matvec-profile: right-hand sides per launch")
    ap.add_argument("--large", action="store_true",
                    help="instead of the headline: the 1024^3 sphere (512 across) matvec rates and memory")
    ap.add_argument("--large-solves", action="store_true", help="with --large: also mixed solves to 1e-5")
    ap.add_argument("--large-solvers", default="bicgstab,qmr")
    ap.add_argument("--large-max-iter", type=int, default=20000)
    ap.add_argument("--orient-dirs", type=int, default=8,
                    help="directions of the extra orientation-batch measurement (0 = skip)")
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="jobs of each seeded Poisson serving stream (0 = skip the serving measurement)")
    ap.add_argument("--serve-burst", type=int, default=48, help="jobs of the capacity burst")
    ap.add_argument("--serve-slots", type=int, default=4, help="worker slots of the serving measurement")
    ap.add_argument("--serve-share", default="0.5,0.25",
                    help="SECONDS,SHARE: the server's solvers yield the GPU past SECONDS (dda_solver_set_gpu_share; "
                         "empty = off)")
    ap.add_argument("--serve-loads", default="0.5,0.9,1.2",
                    help="Poisson arrival rates of the serving streams, as fractions of the burst capacity")
    ap.add_argument("--serve-only", action="store_true",
                    help="only the serving measurement (one JSON line); the full run starts itself this way")
    ap.add_argument("--no-large-grid", action="store_true",
                    help="skip the 1024^3 (sphere 512 across) FP32 / FP64 matvec figures")
    ap.add_argument("--no-large-solves", action="store_true",
                    help="skip the mixed BiCGStab solve to 1e-5 on the 1024^3 grid (about a minute)")
    ap.add_argument("--matvec-profile", type=int, default=0, metavar="K",
                    help="instead of solves: one 512 MiB device copy (counter calibration), then K FP32 and "
                         "K FP64 matvecs; prints per-pass times (for rocprofv3 counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the solvers run on one GPU; --gpus %d is not supported (the multi-GPU slab "
              "decomposition is not implemented)" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import dda_amd as D
    from dda_amd import problem as P
    if args.serve_only:
        print(json.dumps(_plain(serving_bench(args, D))))
        return 0
    if args.large:
        return large_bench(args, D, P, np)

    m = 1.5 + 0.1j
    d = 1e-7
    kd = 0.5 / abs(m)
    lam = 2 * math.pi * d / kd
    g = P.sphere(args.diameter / 2.0, m, d)
    params = P.DDAParams(g, lam)
    nv, S, grid, box = params.nv, params.vec_stride, params.grid, g.box
    rows = np.unique(g.pos[:, 2].astype(np.int64) * box[1] + g.pos[:, 1]).size
    prop, pol1, pol2 = P.beam_pols((0, 0, 1))
    rhs = []
    for pol in (pol1, pol2):
        cc, cs = P.couple_constants(params, prop, pol, "ldr")
        rhs.append((pol, cc, cs, P.einc_plane(params, prop, pol)))

    ctx = D.Context(0)
    t0 = time.perf_counter()
    op = D.GpuMatVec(ctx, g.pos, g.mat, box, d, params.wave_num, precision=64, grid=grid)
    tuning = op.tune(1) if not args.no_tune else "not tuned"  # kernel variants; FP32 twins follow
    t_setup = time.perf_counter() - t0
    modes = [m.strip() for m in args.runs.split(",") if m.strip()]
    if args.matvec_profile:
        return matvec_profile(ctx, op, g, params, rhs, args.matvec_profile, D, P, np, args.profile_rhs)

    def run_mode(run):
        parts = run.split(":")
        mode, kind, batch = parts[0], parts[1], len(parts) > 2 and parts[2] == "batch"
        solver = D.SOLVER_CLASSES[kind](op, max_iter=20000, mixed=(mode == "mixed"))

        # incident fields resident in HBM before the timed region; the
        # polarisations stay there until it ends (the boundary's device path)
        nr, vs = len(rhs), op.vec_size
        einc_dev = D.DeviceVector(ctx, nr * vs, 64)
        einc_dev.upload(np.concatenate([r[3] for r in rhs]))
        pol_dev = D.DeviceVector(ctx, nr * vs, 64)
        one = [(D.DeviceVector(ctx, vs, 64), D.DeviceVector(ctx, vs, 64)) for _ in rhs]
        for (e_d, _), r in zip(one, rhs):
            e_d.upload(r[3])

        def step():
            if batch:
                op.set_couple_constants_batch(np.array([r[2] for r in rhs]))
                infos = [dict(i) for i in solver.solve_device(einc_dev, pol_dev, nr, args.eps)]
            else:
                infos = []
                for (pol, cc, cs, einc), (e_d, p_d) in zip(rhs, one):
                    solver.set_couple_constants(cs)
                    infos.append(dict(solver.solve_device(e_d, p_d, 1, args.eps)[0]))
            bad = [i for i in infos if i["status"] != D.OK]
            if bad:
                raise RuntimeError("solve failed: %s" % bad)
            return infos

        def results(infos):
            ps = pol_dev.download().reshape(nr, vs) if batch else [p_d.download() for _, p_d in one]
            return [(p, inf, cc, einc, pol) for (pol, cc, cs, einc), p, inf in zip(rhs, ps, infos)]

        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        # the timed region: the production path (iterations replayed as HIP
        # graphs), no per-launch events
        t0 = time.perf_counter()
        infos = []
        for _ in range(args.steps):
            infos = step()
        ctx.synchronize()
        el = time.perf_counter() - t0
        res = results(infos)
        # the same steps again with HIP events around every matvec kernel on
        # its stream (eager launches: events cannot sit in a replayed graph):
        # the per-pass times and the roofline come from this second timed region
        op.profile(True)
        op.profile_read()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        ctx.synchronize()
        el_prof = time.perf_counter() - t0
        prof = solver.profile_read()  # the operator the iteration runs on
        op.profile_read()
        op.profile(False)
        # with --host-step, one more step through the host-array path (PCIe
        # copies of the fields and polarisations included)
        el_host = None
        if args.host_step:
            t0 = time.perf_counter()
            if batch:
                op.set_couple_constants_batch(np.array([r[2] for r in rhs]))
                solver.get_pol_vecs(np.array([r[3] for r in rhs]), args.eps)
            else:
                for pol, cc, cs, einc in rhs:
                    solver.set_couple_constants(cs)
                    solver.get_pol_vec(einc, args.eps)
            el_host = time.perf_counter() - t0
        solver.close()
        for v in [einc_dev, pol_dev] + [v for pair in one for v in pair]:
            v.close()
        return res, el, prof, el_host, el_prof

    runs = {}
    for mode in modes:
        runs[mode] = run_mode(mode)
    final_run = modes[-1]
    final_mode, final_solver = final_run.split(":")[:2]
    final_batch = final_run.endswith(":batch")
    results, elapsed, _, elapsed_host, elapsed_prof = runs[final_run]
    ms, cnt, skipped = runs[final_run][2]

    nsolves = 2 * args.steps
    value = nsolves / elapsed
    c = 8 if final_mode == "mixed" else 16
    kb = len(rhs) if final_batch else 1  # right-hand sides per matvec launch
    pb = pass_bytes(nv, S, box, grid, rows, c, kb)
    avg_ms = [ms[i] / cnt[i] if cnt[i] else float("nan") for i in range(5)]
    dom = int(np.nanargmax(ms))
    achieved = pb[dom] / (avg_ms[dom] * 1e-3) / 1e9
    mv_ms = float(np.nansum(avg_ms))
    bmv = b_mv(nv, box, grid, c, kb)

    # observables of the last step, checked against Mie theory (volume-equivalent sphere)
    a_eq = (3 * nv / (4 * math.pi)) ** (1 / 3) * d

    def observe(res):
        out = []
        for p, info, cc, einc, pol in res:
            out.append({"iterations": int(info["iterations"]), "refinements": int(info["refinements"]),
                        "true_rel_residual": info["true_rel_residual"], "seconds": info["seconds"],
                        "cext": P.ext_cross(params, p, einc), "cabs": P.abs_cross(params, p, cc),
                        "cext_forward": P.ext_cross_forward(params, p, prop, pol)})
        return out

    obs = observe(results)
    modes_out = {}
    for mode, (res, el, prof, el_host, el_prof) in runs.items():
        o = observe(res)
        pms, pcnt, _ = prof
        cm = 8 if mode.startswith("mixed") else 16
        km = len(rhs) if mode.endswith(":batch") else 1
        pbm = pass_bytes(nv, S, box, grid, rows, cm, km)
        pav = [pms[i] / pcnt[i] if pcnt[i] else float("nan") for i in range(5)]
        modes_out[mode] = {"solves_per_s": nsolves / el, "ms_per_step": 1e3 * el / args.steps,
                           "ms_per_step_with_events": 1e3 * el_prof / args.steps,
                           "solves_per_s_host_arrays": 2 / el_host if el_host else None,
                           "matvec_ms": float(np.nansum(pav)),
                           "rhs_per_matvec": km,
                           "matvec_gbs_vs_bmv": b_mv(nv, box, grid, cm, km) / (float(np.nansum(pav)) * 1e-3) / 1e9,
                           "pass_avg_us": [1e3 * v for v in pav],
                           "pass_gbs": [pbm[i] / (pav[i] * 1e-3) / 1e9 for i in range(5)],
                           "iterations": [x["iterations"] for x in o],
                           "refinements": [x["refinements"] for x in o],
                           "true_rel_residual": [x["true_rel_residual"] for x in o],
                           "cext": [x["cext"] for x in o], "cabs": [x["cabs"] for x in o]}
    # the CPU oracle's FP64 QMR-CS answer at this workload (to 1e-8), written by
    # tests/golden/make_c2_fixture.py at grid unit 0.5 um; cross sections
    # scale with d^2 at fixed kd and m
    fixture = os.path.join(ROOT, "tests", "golden", "c2_sphere128_fp64.json")
    if args.diameter == 128 and os.path.exists(fixture):
        with open(fixture) as f:
            fx = json.load(f)
        scale = (d / 0.5e-6) ** 2
        ref = [(p["cext"] * scale, p["cabs"] * scale) for p in fx["pols"]]
        for m_ in modes_out.values():
            m_["vs_cpu_oracle"] = {
                "cext_rel_err": max(abs(a / r[0] - 1) for a, r in zip(m_["cext"], ref)),
                "cabs_rel_err": max(abs(a / r[1] - 1) for a, r in zip(m_["cabs"], ref)),
                "reference": "CPU oracle FP64 QMR-CS to 1e-8 (tests/golden/c2_sphere128_fp64.json)"}
    if "fp64:qmr" in modes_out:
        f = modes_out["fp64:qmr"]
        for k in list(modes_out):
            if k != "fp64:qmr":
                m_ = modes_out[k]
                m_["vs_fp64_qmr"] = {
                    "speedup": m_["solves_per_s"] / f["solves_per_s"],
                    "cext_rel_err": max(abs(a / b - 1) for a, b in zip(m_["cext"], f["cext"])),
                    "cabs_rel_err": max(abs(a / b - 1) for a, b in zip(m_["cabs"], f["cabs"]))}
    mie = None
    try:
        import oracle as O
        qext, qsca, qback, gs = O.mie(params.wave_num * a_eq, m)
        area = math.pi * a_eq * a_eq
        mie = {"qext": qext, "qabs": qext - qsca, "cext_rel_err": obs[0]["cext"] / (qext * area) - 1,
               "cabs_rel_err": obs[0]["cabs"] / ((qext - qsca) * area) - 1}
    except Exception as e:  # the oracle is only a checker here
        mie = {"error": str(e)}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            import oracle as O
            dm = op.dmatrix()
            orc = O.Problem(g.pos, g.mat, box, grid, rhs[0][2], d, params.wave_num, dmat=dm)
            del dm
            x = P.random_dipole_vector(params)
            threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
            # MatVecCpu restated over scipy.fft (pocketfft), the stand-in for the
            # reference's FFTW path: serial (the reference's CPU path is
            # single-threaded) and on `threads` threads
            orc.matvec_pocketfft(x, workers=threads)  # warm-up (plans, tensor layout)
            t0 = time.perf_counter()
            orc.matvec_pocketfft(x, workers=1)
            t_serial = time.perf_counter() - t0
            reps = 3
            t0 = time.perf_counter()
            for _ in range(reps):
                orc.matvec_pocketfft(x, workers=threads)
            t_mv = (time.perf_counter() - t0) / reps
            # the oracle's own C++ matvec (recursive FFT) on the same threads, for reference
            O.set_threads(threads)
            t0 = time.perf_counter()
            orc.matvec(x)
            t_orc = time.perf_counter() - t0
            O.set_threads(1)
            fq = modes_out.get("fp64:qmr", {}).get("iterations", [obs[0]["iterations"]])
            mv_per_solve = fq[0] + 2
            model, ncpu = cpu_info()
            cpu = {"value": 1.0 / (t_mv * mv_per_solve), "unit": "solves/s", "cores": threads, "kind": "port",
                   "single_thread_value": 1.0 / (t_serial * mv_per_solve),
                   "oracle_cpp_value": 1.0 / (t_orc * mv_per_solve),
                   "sample": "MatVecCpu restated over scipy.fft (pocketfft, in place of the reference's FFTW; FP64) "
                             "at this workload: %d matvecs on %d threads, %.2f s each (one serial matvec, as the "
                             "reference's single-threaded CPU path: %.2f s; the oracle's C++ matvec with its own "
                             "recursive FFT on %d threads: %.2f s), on %s (%d logical CPUs), scaled by the %d "
                             "matvecs of one FP64 QMR-CS solve; vector updates not counted"
                             % (reps, threads, t_mv, t_serial, threads, t_orc, model, ncpu, mv_per_solve)}
        except Exception as e:
            cpu = {"value": None, "unit": "solves/s", "cores": 1, "kind": "port", "sample": "failed: %s" % e}

    # the reference's single-solve path: a sphere lit along z is rotation
    # symmetric (DDAMain.cpp:105-121, 252-260), so one solve (polarisation 1)
    # gives both polarisations' results (DDAMain.cpp:380-410, 461-464)
    symmetric = None
    if P.rotation_symmetric(g, prop):
        try:
            ss = D.GpuBicgStab(op, max_iter=20000, mixed=True)
            pol_d = D.DeviceVector(ctx, op.vec_size, 64)
            D.solve_particle(ss, params, prop, args.eps, pol_dev=pol_d)  # warm-up
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                r_ = D.solve_particle(ss, params, prop, args.eps, pol_dev=pol_d)
            ctx.synchronize()
            el_s = time.perf_counter() - t0
            base = modes_out.get("fp64:qmr", {}).get("solves_per_s")
            symmetric = {"workload": "the same sphere and direction, one mixed BiCGStab solve per particle "
                                     "(polarisation 1; polarisation 2 by the quarter-turn symmetry)",
                         "particles_per_s": args.steps / el_s,
                         "vs_fp64_qmr_both_polarisations": (args.steps / el_s) / (base / 2) if base else None,
                         "vs_fp64_qmr_one_polarisation": (args.steps / el_s) / base if base else None,
                         "iterations": int(r_["infos"][0]["iterations"]),
                         "cext": [float(c) for c in r_["cext"]], "cabs": [float(c) for c in r_["cabs"]]}
            ss.close()
            pol_d.close()
        except Exception as e:
            symmetric = {"error": str(e)}

    # per-apply matvec times: 3 warm-ups, then 20 applies, each the sum of its
    # five kernels' HIP-event durations (median, p10, p90)
    matvec_applies = None
    if not args.no_matvec_applies:
        try:
            matvec_applies = matvec_leg(ctx, op, g, params, rhs, D, P, np)
        except Exception as e:  # reported, not fatal
            matvec_applies = {"error": str(e)}

    orient = None
    if args.orient_dirs > 0:
        try:
            dirs = P.directions(args.orient_dirs, 2024)
            einc_b, ccs_b, props_b, pols_b = P.incident_batch(params, dirs)
            k = len(ccs_b)
            op.set_couple_constants_batch(ccs_b)
            sb = D.GpuBicgStab(op, max_iter=20000, mixed=True)
            e_d = D.DeviceVector(ctx, k * op.vec_size, 64)
            e_d.upload(einc_b.reshape(-1))
            p_d = D.DeviceVector(ctx, k * op.vec_size, 64)
            sb.solve_device(e_d, p_d, k, args.eps)  # warm-up
            ctx.synchronize()
            t0 = time.perf_counter()
            infos_b = sb.solve_device(e_d, p_d, k, args.eps)
            ctx.synchronize()
            el = time.perf_counter() - t0
            sb.close()
            e_d.close()
            p_d.close()
            base = modes_out.get("fp64:qmr", {}).get("solves_per_s")
            orient = {"workload": "the same sphere, %d seeded directions x 2 polarisations = %d right-hand sides in "
                                  "one block, mixed BiCGStab to %g" % (args.orient_dirs, k, args.eps),
                      "solves_per_s": k / el, "seconds": el,
                      "vs_fp64_qmr": (k / el) / base if base else None,
                      "iterations": [int(i["iterations"]) for i in infos_b],
                      "ok": sum(1 for i in infos_b if i["status"] == D.OK)}
        except Exception as e:
            orient = {"error": str(e)}

    c3 = None
    if not args.no_c3:
        try:
            c3 = c3_bench(args, D, P, np, ctx)
        except Exception as e:  # reported, not fatal
            c3 = {"error": str(e)}

    serving = None
    if args.serve_jobs > 0:
        try:
            serving = serving_child(args)
        except Exception as e:  # reported, not fatal: the headline is the solve rate of the 128-across sphere
            serving = {"error": str(e)}

    large = None
    if not args.no_large_grid:
        try:
            op.close()  # the headline operator's memory back before the 1024^3 operators
            large = large_matvec(args, D, P, np, ctx, solvers=() if args.no_large_solves else ("bicgstab",))
        except Exception as e:  # reported, not fatal
            large = {"error": str(e)}

    info = ctx.info()
    # what a plain device-to-device copy sustains on this GPU (1 GiB, read +
    # write counted): the practical ceiling of a streaming pass
    na = (1 << 30) // 16
    ca, cb = D.DeviceVector(ctx, na, 64), D.DeviceVector(ctx, na, 64)
    cb.copy_from(ca)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        cb.copy_from(ca)
    ctx.synchronize()
    copy_gbs = 2 * (1 << 30) * 10 / (time.perf_counter() - t0) / 1e9
    ca.close()
    cb.close()
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if final_mode == "mixed" else "f64",
        "data": "synthetic (generated sphere geometry, plane-wave incident field)",
        "config": {"workload": "sphere %d dipoles across (nv=%d), m=1.5+0.1i, |m|kd=0.5, grid %s, prop z, "
                               "2 polarisations%s, %s to resid %g, %s" % (
                                   args.diameter, nv, "x".join(map(str, grid)),
                                   " solved as one 2-RHS block" if final_batch else " solved one after the other",
                                   final_solver, args.eps,
                                   "FP32 Krylov in FP64 refinement" if final_mode == "mixed" else "FP64"),
                   "nv": nv, "grid": list(grid), "solver": final_solver, "mode": final_mode, "batch": final_batch,
                   "eps": args.eps,
                   "baseline": "fp64:qmr (the reference's default solver, all FP64)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "copy_gbs": copy_gbs,
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.diameter, grid, "fp32" if c == 8 else "fp64", dom, kb),
                     "traffic_source": "profiles/traffic_matvec_sphere128.md (rocprofv3 FETCH_SIZE/WRITE_SIZE, "
                                       "calibrated)",
                     "kernel": PASS_NAMES[dom],
                     "bytes_per_launch": pb[dom], "avg_launch_us": 1e3 * avg_ms[dom],
                     "launches": int(cnt[dom]), "noop_launches_excluded": skipped,
                     "measured_in": "second timed region of the %s run, HIP events around every matvec kernel "
                                    "(%s kernels; %.1f ms per step there vs %.1f ms with graphs and no events)" % (
                                        final_run, "FP32" if c == 8 else "FP64", 1e3 * elapsed_prof / args.steps,
                                        1e3 * elapsed / args.steps)},
        "matvec": {"ms": mv_ms, "rhs_per_launch": kb, "bmv_bytes": bmv, "gbs_vs_bmv": bmv / (mv_ms * 1e-3) / 1e9,
                   "frac_vs_bmv": bmv / (mv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                   "pass_avg_us": [1e3 * v for v in avg_ms], "pass_bytes": pb,
                   "pass_gbs": [pb[i] / (avg_ms[i] * 1e-3) / 1e9 for i in range(5)]},
        "solves": obs,
        "modes": modes_out,
        "mie": mie,
        "setup_s": t_setup,
        "tuning": tuning,
        "device": info,
        "cpu_baseline": cpu,
        "symmetric_single_solve": symmetric,
        "matvec_applies": matvec_applies,
        "orientation_batch": orient,
        "c3_batch": c3,
        "serving": serving,
        "large_grid": large,
    }
    print(json.dumps(_plain(line)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
