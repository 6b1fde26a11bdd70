"""Writes tests/golden/c2_sphere128_fp64.json: the CPU oracle's FP64 answer for
the sphere 128 dipoles across (1,099,136 dipoles, m = 1.5 + 0.1i,
|m| kd = 0.5, grid 256^3), plane wave along z, both polarisations, LDR,
QMR-CS (DDA/QmrCs.cpp restated) to relative residual 1e-8: Cext (Beam.cpp:58-68),
Cabs (AbsCross.cpp:27-48), the forward-amplitude Cext (Beam.cpp:125-133) and
the iteration counts. The GPU tests compare the device solves against it.

Run in this container (about half an hour on 8 cores):
    python tests/golden/make_c2_fixture.py
"""
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from dda_amd import problem as P  # noqa: E402


def main():
    m, d, eps = 1.5 + 0.1j, 0.5e-6, 1e-8
    lam = 2 * math.pi * d * abs(m) / 0.5
    g = P.sphere(64.0, m, d)
    params = P.DDAParams(g, lam)
    assert params.nv == 1099136 and tuple(params.grid) == (256, 256, 256)
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1))
    t0 = time.time()
    dmat = O.dmatrix(g.box, params.grid, d, params.wave_num, precision=64)
    print("tensor %.0f s" % (time.time() - t0), flush=True)
    prop, pol1, pol2 = P.beam_pols((0, 0, 1))
    out = {"workload": "sphere 128 dipoles across, m=1.5+0.1i, |m|kd=0.5, d=0.5um, grid 256^3, prop +z, LDR",
           "nv": params.nv, "grid": list(params.grid), "solver": "qmr", "eps": eps, "pols": []}
    for pol in (pol1, pol2):
        cc, cs = P.couple_constants(params, prop, pol, "ldr")
        einc = P.einc_plane(params, prop, pol)
        orc = O.Problem(g.pos, g.mat, g.box, params.grid, cs, d, params.wave_num, dmat=dmat)
        t0 = time.time()
        st, p, it, hist = orc.solve("qmr", einc, eps, max_iter=5000)
        assert st == O.ST_OK, st
        out["pols"].append({"pol": list(map(float, pol)), "iterations": int(it),
                            "cext": orc.ext_cross_sum(p, einc), "cabs": orc.abs_cross(p, cc),
                            "cext_forward": orc.ext_cross_forward(p, prop, pol),
                            "seconds": time.time() - t0})
        print(out["pols"][-1], flush=True)
    with open(os.path.join(HERE, "c2_sphere128_fp64.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
