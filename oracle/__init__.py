"""ctypes bindings for the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: the checker for the HIP path. Only the test suite, the
smoke check and the CPU-baseline leg of the benchmark may import this module; the
product package never does. Each wrapper names the reference file it restates
(see the header of oracle.cpp for what pins each function).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

ST_OK, ST_INVALID, ST_BREAKDOWN, ST_MAXITER, ST_STAGNATION = 0, 1, 4, 5, 6
SOLVERS = {"qmr": 0, "bicgstab": 1, "cgnr": 2, "bicg": 3}
SHAPES = {"sphere": 0, "box": 1, "cylinder": 2, "ellipsoid": 3}

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class _Prob(ctypes.Structure):
    _fields_ = [("box", ctypes.c_uint32 * 3), ("grid", ctypes.c_uint32 * 3),
                ("nv", ctypes.c_uint32), ("vec_stride", ctypes.c_uint32),
                ("pos", _u32p), ("mat", _u8p), ("nmat", ctypes.c_uint32),
                ("cc_sqrt", _dp), ("dmatrix", _dp)]


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_fft.argtypes = [ctypes.c_int, ctypes.c_long, _dp, ctypes.c_int]
        L.orc_fft_fit.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.orc_fft_fit.restype = ctypes.c_uint32
        L.orc_geometry.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u32p, _u32p]
        L.orc_geometry.restype = ctypes.c_long
        L.orc_beam_pols.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_cc.argtypes = [ctypes.c_int, ctypes.c_uint32, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_cc_planted.argtypes = [ctypes.c_int, ctypes.c_uint32, _dp, ctypes.c_double, ctypes.c_double, _dp, _dp,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_einc_plane.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, _dp]
        L.orc_dmatrix.argtypes = [_u32p, _u32p, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_dmatrix64.argtypes = [_u32p, _u32p, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_green.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_matvec.argtypes = [ctypes.POINTER(_Prob), _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_matvec_direct.argtypes = [ctypes.POINTER(_Prob), ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.c_int]
        L.orc_solve.argtypes = [ctypes.POINTER(_Prob), ctypes.c_int, _dp, ctypes.c_double, ctypes.c_uint64, _dp,
                                _dp, _u64p, _dp, ctypes.c_uint64]
        L.orc_solve_ex.argtypes = [ctypes.POINTER(_Prob), ctypes.c_int, _dp, ctypes.c_double, ctypes.c_uint64, _dp,
                                   _dp, _u64p, _dp, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_ext_cross_sum.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _dp, ctypes.c_double]
        L.orc_ext_cross_sum.restype = ctypes.c_double
        L.orc_abs_cross.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _u8p, ctypes.c_uint32, _dp, ctypes.c_double]
        L.orc_abs_cross.restype = ctypes.c_double
        L.orc_calc_field.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _u32p, _u32p, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp]
        L.orc_ext_cross_forward.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _u32p, _u32p, ctypes.c_double,
                                            ctypes.c_double, _dp, _dp]
        L.orc_ext_cross_forward.restype = ctypes.c_double
        L.orc_jones.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _dp, _dp, ctypes.c_int, _u32p, _u32p,
                                ctypes.c_double, ctypes.c_double, _dp, _dp, _dp,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                ctypes.c_uint32, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_mie.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _cd(a):
    """complex128 array -> contiguous complex128 (viewed as interleaved doubles)."""
    return np.ascontiguousarray(a, dtype=np.complex128)


def set_threads(n):
    """OpenMP threads of the FFT matvec (1 = serial, as the reference's CPU path)."""
    lib().orc_set_threads(int(n))


def fft(data, inverse=False):
    """Batched unnormalised FFT along the last axis (FFTWPlan.cpp:82,113)."""
    d = _cd(data).copy()
    n = d.shape[-1]
    lib().orc_fft(n, d.size // n, _p(d.view(np.float64), _dp), int(inverse))
    return d


def fft_fit(x, override=0, support_non_pot=True):
    """DDAParams.cpp:45-90."""
    return int(lib().orc_fft_fit(x, override, int(support_non_pot)))


def geometry(shape, p0, p1=0.0, p2=0.0):
    """Shapes.cpp:53-143 (+ a per-axis ellipsoid). Returns (pos [nv,3] uint32, box)."""
    s = SHAPES[shape]
    box = np.zeros(3, np.uint32)
    n = lib().orc_geometry(s, p0, p1, p2, None, _p(box, _u32p))
    pos = np.zeros((max(n, 0), 3), np.uint32)
    lib().orc_geometry(s, p0, p1, p2, _p(pos, _u32p), _p(box, _u32p))
    return pos, box


def beam_pols(prop):
    """Beam.cpp:34-51 -> (prop, pol1 'y', pol2 'x')."""
    pin = np.asarray(prop, np.float64)
    out = [np.zeros(3) for _ in range(3)]
    lib().orc_beam_pols(_p(pin, _dp), *[_p(o, _dp) for o in out])
    return out


def couple_constants(kind, m, grid_unit, kd, prop, incpol, planted=None):
    """PolarizabilityDescription.cpp:65-110 -> (cc, cc_sqrt) as [nmat,3] complex.
    planted = (b1, b2, b3, s_factor) replaces the LDR constants and scales S: a
    mistake planted by a test to show that its check rejects it."""
    kinds = {"cm": 0, "ldr": 1, "ldr-avgpol": 2}
    m = np.asarray(m, np.complex128)
    if m.ndim == 0:
        m = np.full((1, 3), m)
    elif m.ndim == 1:
        m = np.repeat(m[:, None], 3, axis=1)
    m = _cd(m)
    cc = np.zeros_like(m)
    cs = np.zeros_like(m)
    pr = np.asarray(prop, np.float64)
    ip = np.asarray(incpol, np.float64)
    if planted is None:
        lib().orc_cc(kinds[kind], m.shape[0], _p(m.view(np.float64), _dp), grid_unit, kd,
                     _p(pr, _dp), _p(ip, _dp), _p(cc.view(np.float64), _dp), _p(cs.view(np.float64), _dp))
    else:
        lib().orc_cc_planted(kinds[kind], m.shape[0], _p(m.view(np.float64), _dp), grid_unit, kd,
                             _p(pr, _dp), _p(ip, _dp), *[float(v) for v in planted],
                             _p(cc.view(np.float64), _dp), _p(cs.view(np.float64), _dp))
    return cc, cs


def einc_plane(pos, box, vec_stride, grid_unit, wave_num, prop, pol):
    """Beam.cpp:110-118 -> SoA [3*vec_stride] complex."""
    pos = np.ascontiguousarray(pos, np.uint32)
    box = np.ascontiguousarray(box, np.uint32)
    out = np.zeros(3 * vec_stride, np.complex128)
    pr = np.asarray(prop, np.float64)
    pl = np.asarray(pol, np.float64)
    lib().orc_einc_plane(pos.shape[0], vec_stride, _p(pos, _u32p), _p(box, _u32p), grid_unit, wave_num,
                         _p(pr, _dp), _p(pl, _dp), _p(out.view(np.float64), _dp))
    return out


def dmatrix(box, grid, grid_unit, wave_num, precision=80):
    """DMatrixCpu.cpp:133-205; returns complex array [Gx, Gz, Gy, 6] (reference memory order).
    precision 80 computes in long double, 64 in double."""
    box = np.ascontiguousarray(box, np.uint32)
    grid = np.ascontiguousarray(grid, np.uint32)
    out = np.zeros((int(grid[0]), int(grid[2]), int(grid[1]), 6), np.complex128)
    f = lib().orc_dmatrix if precision == 80 else lib().orc_dmatrix64
    st = f(_p(box, _u32p), _p(grid, _u32p), grid_unit, wave_num, _p(out.view(np.float64), _dp))
    if st != ST_OK:
        raise ValueError("grid too small for box")
    return out


def green(r, wave_num, mu, nu):
    out = np.zeros(2)
    lib().orc_green(float(r[0]), float(r[1]), float(r[2]), wave_num, mu, nu, _p(out, _dp))
    return complex(out[0], out[1])


class Problem:
    """Geometry + grid + couple constants + D matrix, as the CPU path holds them."""

    def __init__(self, pos, mat, box, grid, cc_sqrt, grid_unit, wave_num, dmat=None, vec_stride=None,
                 dmat_precision=80):
        self.pos = np.ascontiguousarray(pos, np.uint32)
        self.nv = self.pos.shape[0]
        self.mat = np.ascontiguousarray(mat if mat is not None else np.zeros(self.nv), np.uint8)
        self.box = np.ascontiguousarray(box, np.uint32)
        self.grid = np.ascontiguousarray(grid, np.uint32)
        self.cc_sqrt = _cd(np.asarray(cc_sqrt).reshape(-1, 3))
        self.grid_unit = float(grid_unit)
        self.wave_num = float(wave_num)
        self.vec_stride = int(vec_stride) if vec_stride else ((self.nv + 15) // 16) * 16
        self.dmat = _cd(dmat) if dmat is not None else dmatrix(self.box, self.grid, self.grid_unit, self.wave_num,
                                                                 dmat_precision)
        self._s = _Prob()
        self._s.box[:] = [int(v) for v in self.box]
        self._s.grid[:] = [int(v) for v in self.grid]
        self._s.nv = self.nv
        self._s.vec_stride = self.vec_stride
        self._s.pos = _p(self.pos, _u32p)
        self._s.mat = _p(self.mat, _u8p)
        self._s.nmat = self.cc_sqrt.shape[0]
        self._s.cc_sqrt = _p(self.cc_sqrt.view(np.float64), _dp)
        self._s.dmatrix = _p(self.dmat.view(np.float64), _dp)

    def matvec(self, x, conj=False, precision=64):
        """MatVecCpu.cpp:67-150."""
        x = _cd(x)
        y = np.zeros(3 * self.vec_stride, np.complex128)
        lib().orc_matvec(ctypes.byref(self._s), _p(x.view(np.float64), _dp), _p(y.view(np.float64), _dp),
                         int(conj), precision)
        return y

    def matvec_pocketfft(self, x, conj=False, workers=1):
        """MatVecCpu.cpp:67-150 restated over numpy arrays with scipy.fft
        (pocketfft, C, `workers` threads) in place of FFTW: the CPU baseline's
        stand-in for the reference's FFTW path (the oracle's own recursive FFT is
        several times slower). Same steps and zero skipping as the reference: X
        FFTs over the box's (y, z) rows, Z FFTs over y < By, Y FFTs over every
        z, the symmetric tensor product, the inverses, the gather. Checked
        against matvec() in the CPU tests."""
        from concurrent.futures import ThreadPoolExecutor
        import scipy.fft as sf
        Gx, Gy, Gz = (int(v) for v in self.grid)
        Bx, By, Bz = (int(v) for v in self.box)
        S, nv = self.vec_stride, self.nv
        if not hasattr(self, "_dT"):  # the tensor as [6][Gz][Gy][Gx] (set-up, not per matvec)
            self._dT = np.ascontiguousarray(np.moveaxis(self.dmat.reshape(Gx, Gz, Gy, 6), (3, 1, 2, 0), (0, 1, 2, 3)))
        xin = _cd(x).reshape(3, S)[:, :nv]
        if conj:
            xin = np.conj(xin)
        px, py, pz = (self.pos[:, a].astype(np.int64) for a in range(3))
        ccs = self.cc_sqrt[self.mat]  # [nv][3]
        X = np.zeros((3, Bz, By, Gx), np.complex128)
        for c in range(3):
            X[c, pz, py, px] = ccs[:, c] * xin[c]
        X = sf.fft(X, axis=3, workers=workers, overwrite_x=True)
        A = np.zeros((3, Gz, Gy, Gx), np.complex128)
        A[:, :Bz, :By, :] = X
        A[:, :, :By, :] = sf.fft(A[:, :, :By, :], axis=1, workers=workers)
        A = sf.fft(A, axis=2, workers=workers, overwrite_x=True)
        D = self._dT

        def product(z0, z1):
            a0, a1, a2 = A[0, z0:z1], A[1, z0:z1], A[2, z0:z1]
            d = [D[e, z0:z1] for e in range(6)]
            y0 = d[0] * a0 + d[1] * a1 + d[2] * a2
            y1 = d[1] * a0 + d[3] * a1 + d[4] * a2
            y2 = d[2] * a0 + d[4] * a1 + d[5] * a2
            A[0, z0:z1], A[1, z0:z1], A[2, z0:z1] = y0, y1, y2

        step = max(1, Gz // (4 * max(1, workers)))
        with ThreadPoolExecutor(max(1, workers)) as ex:
            list(ex.map(lambda z: product(z, min(Gz, z + step)), range(0, Gz, step)))
        A = sf.ifft(A, axis=2, workers=workers, norm="forward", overwrite_x=True)
        A[:, :, :By, :] = sf.ifft(A[:, :, :By, :], axis=1, workers=workers, norm="forward")
        X = sf.ifft(np.ascontiguousarray(A[:, :Bz, :By, :]), axis=3, workers=workers, norm="forward",
                    overwrite_x=True)
        y = np.zeros((3, S), np.complex128)
        for c in range(3):
            y[c, :nv] = ccs[:, c] * X[c, pz, py, px] + xin[c]
        if conj:
            y = np.conj(y)
        return y.reshape(-1)

    def matvec_direct(self, x, conj=False):
        """Dense O(nv^2) direct sum in long double."""
        x = _cd(x)
        y = np.zeros(3 * self.vec_stride, np.complex128)
        lib().orc_matvec_direct(ctypes.byref(self._s), self.grid_unit, self.wave_num,
                                _p(x.view(np.float64), _dp), _p(y.view(np.float64), _dp), int(conj))
        return y

    def solve(self, kind, einc, eps, max_iter=0, start=None, max_iter_sample=0, max_res_increase=0):
        """IterativeSolverBase.cpp:40-103 + the CPU solver `kind`.
        max_res_increase = 0 takes the solver's default.
        Returns (status, pvec, iterations, residual history)."""
        einc = _cd(einc)
        pvec = np.zeros(3 * self.vec_stride, np.complex128)
        iters = ctypes.c_uint64(0)
        mi = max_iter if max_iter else 3 * int(np.prod(self.box.astype(np.int64)))
        hist = np.full(min(mi, 100000) + 1, np.nan)
        st = None if start is None else _cd(start)
        status = lib().orc_solve_ex(ctypes.byref(self._s), SOLVERS[kind], _p(einc.view(np.float64), _dp), eps,
                                    max_iter, None if st is None else _p(st.view(np.float64), _dp),
                                    _p(pvec.view(np.float64), _dp), ctypes.byref(iters),
                                    _p(hist, _dp) if mi <= 100000 else None, max_iter_sample, max_res_increase)
        n = int(iters.value)
        return status, pvec, n, hist[: n + 1]

    def ext_cross_sum(self, pvec, einc):
        """Beam.cpp:58-68."""
        return lib().orc_ext_cross_sum(self.nv, self.vec_stride, _p(_cd(pvec).view(np.float64), _dp),
                                       _p(_cd(einc).view(np.float64), _dp), self.wave_num)

    def abs_cross(self, pvec, cc):
        """AbsCross.cpp:27-48."""
        cc = _cd(np.asarray(cc).reshape(-1, 3))
        return lib().orc_abs_cross(self.nv, self.vec_stride, _p(_cd(pvec).view(np.float64), _dp),
                                   _p(self.mat, _u8p), cc.shape[0], _p(cc.view(np.float64), _dp), self.wave_num)

    def calc_field(self, pvec, n):
        """CpuFieldCalculator.cpp:48-98."""
        out = np.zeros(3, np.complex128)
        nn = np.asarray(n, np.float64)
        lib().orc_calc_field(self.nv, self.vec_stride, _p(_cd(pvec).view(np.float64), _dp), _p(self.pos, _u32p),
                             _p(self.box, _u32p), self.grid_unit, self.wave_num, _p(nn, _dp),
                             _p(out.view(np.float64), _dp))
        return out

    def ext_cross_forward(self, pvec, prop, pol):
        """Beam.cpp:125-133."""
        pr = np.asarray(prop, np.float64)
        pl = np.asarray(pol, np.float64)
        return lib().orc_ext_cross_forward(self.nv, self.vec_stride, _p(_cd(pvec).view(np.float64), _dp),
                                           _p(self.pos, _u32p), _p(self.box, _u32p), self.grid_unit,
                                           self.wave_num, _p(pr, _dp), _p(pl, _dp))

    def jones(self, p1, p2, symmetric, prop, pol1, pol2, ntheta, theta0, dtheta, nphi, phi0, dphi):
        """FarFieldCalc.cpp:119-174 + JonesCalculus.cpp:74-86 -> [count,2,2] complex."""
        count = ntheta * nphi
        out = np.zeros((count, 2, 2), np.complex128)
        args = [np.asarray(v, np.float64) for v in (prop, pol1, pol2)]
        lib().orc_jones(self.nv, self.vec_stride, _p(_cd(p1).view(np.float64), _dp),
                        _p(_cd(p2 if p2 is not None else p1).view(np.float64), _dp), int(symmetric),
                        _p(self.pos, _u32p), _p(self.box, _u32p), self.grid_unit, self.wave_num,
                        *[_p(a, _dp) for a in args], ntheta, theta0, dtheta, nphi, phi0, dphi,
                        _p(out.view(np.float64), _dp))
        return out


def mie(x, m):
    """EMSim/Mie.cpp:40-174 -> (qext, qsca, qback, g)."""
    outs = [ctypes.c_double() for _ in range(4)]
    lib().orc_mie(x, complex(m).real, complex(m).imag, *[ctypes.byref(o) for o in outs])
    return tuple(o.value for o in outs)
