"""Scattering-job streams for the job server (dda_amd.Server).

A job is one particle with two incident polarisations. The stream follows the
serving workload: Poisson arrivals (exponential gaps), shape drawn uniformly
from {sphere, box, ellipsoid}, largest extent uniform in 16..96 dipoles with the
other axes scaled by U(0.5, 1), Re m ~ U(1.2, 2.0), Im m ~ U(0, 0.2),
kd = 0.5 / |m|, a uniformly random orientation (the incident frame rotated by a
random unit quaternion) and both polarisations of Beam.cpp:34-51 in that
frame, LDR polarisabilities. Everything is drawn from numpy's PCG64 with the
given seed.
"""
import math
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import problem as P

GRID_UNIT = 1e-7


@dataclass
class Job:
    shape: str
    dims: tuple
    m: complex
    geometry: object
    params: object
    prop: np.ndarray
    pols: np.ndarray
    einc: np.ndarray          # [2, 3 * vec_stride]
    cc: np.ndarray            # [2, nmat, 3]
    cc_sqrt: np.ndarray       # [2, nmat, 3]
    arrival: float = 0.0      # seconds after the stream starts
    result: dict = field(default_factory=dict)


def _rotation(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def make_job(rng, shape=None, dim=None, m=None, other=None):
    shape = shape or ["sphere", "box", "ellipsoid"][rng.integers(3)]
    dim = int(dim or rng.integers(16, 97))
    other = list(other) if other is not None else [max(1, int(round(dim * rng.uniform(0.5, 1.0)))) for _ in range(2)]
    m = m if m is not None else complex(rng.uniform(1.2, 2.0), rng.uniform(0.0, 0.2))
    if shape == "sphere":
        g = P.sphere(dim / 2.0, m, GRID_UNIT)
        dims = (dim, dim, dim)
    elif shape == "box":
        g = P.box(dim, other[0], other[1], m, GRID_UNIT)
        dims = (dim, other[0], other[1])
    else:
        g = P.ellipsoid(dim / 2.0, other[0] / 2.0, other[1] / 2.0, m, GRID_UNIT)
        dims = (dim, other[0], other[1])
    kd = 0.5 / abs(m)
    params = P.DDAParams(g, 2 * math.pi * GRID_UNIT / kd)
    rot = _rotation(rng.normal(size=4))
    prop = rot @ np.array([0.0, 0.0, 1.0])
    _, pol1, pol2 = P.beam_pols((0.0, 0.0, 1.0))
    pols = np.array([rot @ pol1, rot @ pol2])
    einc = np.array([P.einc_plane(params, prop, p) for p in pols])
    ccs = [P.couple_constants(params, prop, p, "ldr") for p in pols]
    return Job(shape, dims, m, g, params, prop, pols, einc, np.array([c for c, _ in ccs]),
               np.array([s for _, s in ccs]))


def make_stream(n, seed, rate=None, label=None):
    """n jobs; with `rate` (jobs/s) Poisson arrival times, else all at t = 0.
    `label`: a progress line on stderr every 500 jobs made."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        jobs.append(make_job(rng))
        if label and (k + 1) % 500 == 0:
            print("serving %s: %d of %d jobs made" % (label, k + 1, n), file=sys.stderr, flush=True)
    t = 0.0
    for j in jobs:
        if rate:
            t += rng.exponential(1.0 / rate)
        j.arrival = t
    return jobs


def submit(server, job, grid=(0, 0, 0)):
    """Submits a job; grid 0 lets the server pick per axis the smallest length
    with compile-time planned kernels (dda_fft_fit_static)."""
    g = job.geometry
    return server.submit(g.pos, g.mat, g.box, g.grid_unit, job.params.wave_num, job.einc, job.cc_sqrt,
                         nmat=len(g.materials), grid=grid)


def run_stream(server, jobs, rate=None, keep_pol=True, label=None):
    """Submits the jobs at their arrival times (re-drawn for `rate` if given,
    keeping the order), waits for all and returns latency and throughput
    figures: latency = completion - submission on the server's clock.
    keep_pol=False drops each job's polarisations once it is done (long
    streams); `label`: a progress line on stderr every 250 jobs submitted and
    every 250 jobs done."""
    if rate:
        rng = np.random.default_rng(12345)
        t = 0.0
        for j in jobs:
            t += rng.exponential(1.0 / rate)
            j.arrival = t
    t0 = time.perf_counter()
    ids = []
    for j in jobs:
        while True:
            dt = j.arrival - (time.perf_counter() - t0)
            if dt <= 0:
                break
            time.sleep(min(dt, 0.002))
        ids.append(submit(server, j))
        if label and len(ids) % 250 == 0:
            print("serving %s: %d of %d jobs submitted" % (label, len(ids), len(jobs)), file=sys.stderr, flush=True)
    for k, (j, i) in enumerate(zip(jobs, ids)):
        pol, infos, res = server.wait(i)
        j.result = {"pol": pol if keep_pol else None, "infos": infos, **res}
        if label and (k + 1) % 250 == 0:
            print("serving %s: %d of %d jobs done" % (label, k + 1, len(jobs)), file=sys.stderr, flush=True)
    lat = np.array([j.result["t_end"] - j.result["t_submit"] for j in jobs])
    wait = np.array([j.result["t_start"] - j.result["t_submit"] for j in jobs])
    serve = np.array([j.result["t_end"] - j.result["t_start"] for j in jobs])
    slow = jobs[int(np.argmax(serve))]
    start = min(j.result["t_submit"] for j in jobs)
    end = max(j.result["t_end"] for j in jobs)
    solves = sum(len(j.result["infos"]) for j in jobs)
    ok = sum(1 for j in jobs if j.result["status"] == 0)
    return {"jobs": len(jobs), "ok": ok, "solves": solves, "makespan_s": end - start,
            "solves_per_s": solves / (end - start), "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "mean_ms": 1e3 * float(lat.mean()),
            "packed": sum(1 for j in jobs if j.result["packed_rhs"] > len(j.result["infos"])),
            "cache_hits": sum(j.result["cache_hit"] for j in jobs),
            "cache_hit_rate": sum(j.result["cache_hit"] for j in jobs) / len(jobs),
            # per job: set-up (operator, tensor, solver) and, of it, the tensor lookup or generation
            "setup_ms_mean": 1e3 * float(np.mean([j.result["setup_s"] for j in jobs])),
            "tensor_ms_mean": 1e3 * float(np.mean([j.result["tensor_s"] for j in jobs])),
            "tensor_ms_p50": 1e3 * float(np.percentile([j.result["tensor_s"] for j in jobs], 50)),
            "tensor_ms_max": 1e3 * float(np.max([j.result["tensor_s"] for j in jobs])),
            # queueing (submission to a slot taking the job) and service (slot to completion)
            "wait_ms_p50": 1e3 * float(np.percentile(wait, 50)), "wait_ms_p99": 1e3 * float(np.percentile(wait, 99)),
            "service_ms_p50": 1e3 * float(np.percentile(serve, 50)),
            "service_ms_p99": 1e3 * float(np.percentile(serve, 99)), "service_ms_max": 1e3 * float(serve.max()),
            "slowest_job": {"box": [int(b) for b in slow.geometry.box], "nv": int(slow.geometry.nv),
                            "m": [float(np.real(slow.m)), float(np.imag(slow.m))],
                            "iterations": int(slow.result["iterations"])}}
