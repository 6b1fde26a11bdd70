"""The C++ adapters (dda_amd/host/dda_hip.hpp) used as the reference's callers
use GpuMatVec / MatVec / GpuQmrCs, checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from dda_amd import problem as P

import cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "dda_amd", "_build", "adapter_check")


def test_cpp_adapters_match_oracle(tmp_path):
    assert os.path.exists(BIN), "build the adapter driver first (__graft_entry__.build())"
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    einc = P.einc_plane(params, prop, pol)
    x = P.random_dipole_vector(params, seed=21)
    eps = 1e-10
    case = tmp_path / "case.bin"
    with open(case, "wb") as f:
        np.array([params.nv, len(g.materials), *g.box, *params.grid], np.uint32).tofile(f)
        np.array([g.grid_unit, params.wave_num, eps], np.float64).tofile(f)
        np.ascontiguousarray(g.pos, np.uint32).tofile(f)
        np.ascontiguousarray(g.mat, np.uint8).tofile(f)
        np.ascontiguousarray(cs, np.complex128).tofile(f)
        x.astype(np.complex128).tofile(f)
        einc.astype(np.complex128).tofile(f)
    out = tmp_path / "out.bin"
    r = subprocess.run([BIN, str(case), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "adapter_check ok" in r.stdout and "batch ok" in r.stdout
    assert r.stdout.count("profiling run:") == 2  # the float and the double solver
    n = params.vec_size
    data = np.fromfile(out, np.complex128, 4 * n)
    it, it32 = (int(v) for v in np.fromfile(out, np.uint64, 2, offset=4 * n * 16))
    y64, y32, p, p32 = data[:n], data[n:2 * n], data[2 * n:3 * n], data[3 * n:]
    orc = cases.oracle_problem(g, params, cs)
    y_ref = orc.matvec(x, precision=80)
    assert cases.rel_err(y64, y_ref) <= 1e-12
    assert cases.rel_err(y32, y_ref) <= 1e-5
    st, p_ref, it_ref, _ = orc.solve("qmr", einc, eps, max_iter=10000)
    assert st == 0 and abs(it - it_ref) <= max(1, it_ref // 10)
    assert cases.rel_err(p, p_ref) <= 1e-8
    # the float path (GpuQmrCs<float>, eps 1e-5): DipPol and the Jones matrices
    # within the reference's float threshold of FieldDiff (DDA/OMakefile:66-67);
    # the two-material sphere along z is quarter-turn symmetric, so its Jones
    # matrices come from the polarisation-1 solve (FarFieldCalc.cpp:167-168)
    from dda_amd import output as OUT
    assert it32 > 0 and OUT.field_diff(p32, p_ref) < 1e-3
    _, pol1, pol2 = P.beam_pols(prop)
    args = (prop, pol1, pol2, 31, 0.0, np.radians(6.0), 20, 0.0, np.radians(18.0))
    dj = OUT.field_diff(orc.jones(p32, None, 1, *args), orc.jones(p_ref, None, 1, *args))
    print("float path: %d iterations, FieldDiff DipPol %.2e, Jones %.2e" % (it32, OUT.field_diff(p32, p_ref), dj))
    assert dj < 1e-3
