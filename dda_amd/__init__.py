"""MI355X-native DDA matvec and Krylov solvers: Python bindings of libdda_hip.

The product is the C ABI in include/dda_hip.h, implemented by hand-written
gfx950 HIP kernels in dda_amd/csrc and built into dda_amd/_lib/libdda_hip.so.
This module binds it with ctypes and mirrors the reference's class interfaces:

  GpuMatVec        DDA/GpuMatVec.hpp:85-99   (device vectors)
  MatVecGpu        DDA/MatVecGpu.cpp:51-68   (host vectors; apply(arg, result, conj))
  GpuQmrCs, GpuBicgStab, GpuCgnr, GpuBicgCs  with getPolVec(einc, eps, start)
                   DDA/IterativeSolverBase.cpp:40-103

There is no CPU fallback: if the library is missing, or no GPU is present,
every entry point raises.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libdda_hip.so")

OK, INVALID_ARG, OUT_OF_MEMORY, HIP_ERROR, BREAKDOWN, MAX_ITER, STAGNATION, NOT_SUPPORTED = range(8)
STATUS_NAMES = {0: "OK", 1: "INVALID_ARG", 2: "OUT_OF_MEMORY", 3: "HIP_ERROR", 4: "BREAKDOWN",
                5: "MAX_ITER", 6: "STAGNATION", 7: "NOT_SUPPORTED"}
QMR_CS, BICGSTAB, CGNR, BICG_CS = 0, 1, 2, 3
MODE_FP64, MODE_MIXED, MODE_FP32 = 0, 1, 2
MODES = {"fp64": MODE_FP64, "mixed": MODE_MIXED, "fp32": MODE_FP32}
SOLVER_KINDS = {"qmr": QMR_CS, "bicgstab": BICGSTAB, "cgnr": CGNR, "bicg": BICG_CS}

# Every symbol include/dda_hip.h declares.
EXPORTS = [
    "dda_last_error", "dda_ctx_create", "dda_ctx_destroy", "dda_ctx_synchronize", "dda_ctx_info",
    "dda_malloc", "dda_free", "dda_memcpy_h2d", "dda_memcpy_d2h", "dda_memcpy_d2d", "dda_fft_fit",
    "dda_fft_fit_static", "dda_grid_plan",
    "dda_op_create", "dda_op_destroy", "dda_op_info", "dda_op_set_cc_sqrt", "dda_op_set_cc_sqrt_batch",
    "dda_op_apply", "dda_op_apply_batch", "dda_op_tune", "dda_op_get_variants", "dda_op_set_variants",
    "dda_op_apply_host", "dda_op_get_dmatrix", "dda_op_profile", "dda_op_profile_read",
    "dda_solver_create", "dda_solver_create_ex", "dda_solver_destroy", "dda_solver_solve", "dda_solver_solve_batch",
    "dda_solver_history",
    "dda_solver_set_inner_floor", "dda_solver_set_inner_plateau", "dda_solver_set_gpu_share", "dda_solver_profile_read",
    "dda_solver_profiling_run",
    "dda_dhat_cache_create", "dda_dhat_cache_destroy", "dda_dhat_cache_stats", "dda_op_create_cached",
    "dda_server_create", "dda_server_submit", "dda_server_wait", "dda_server_stats", "dda_server_destroy",
    "dda_solver_solve_batch_dev", "dda_couple_constants", "dda_couple_constants_diag", "dda_incident_plane",
    "dda_cross_sections",
    "dda_cross_sections_noalloc",
    "dda_far_field", "dda_jones", "dda_op_create_pack", "dda_op_pack_info", "dda_op_pack_map",
    "dda_op_set_rhs_members",
]
POL_CM, POL_LDR, POL_LDR_AVGPOL = 0, 1, 2
POL_KINDS = {"cm": POL_CM, "ldr": POL_LDR, "ldr-avgpol": POL_LDR_AVGPOL}


class DDAError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class SolverError(DDAError):
    """Raised for BREAKDOWN / MAX_ITER / STAGNATION (the reference's ABORT_MSG/ASSERT)."""


class OpDesc(ctypes.Structure):
    _fields_ = [("box", ctypes.c_uint32 * 3), ("grid", ctypes.c_uint32 * 3), ("nv", ctypes.c_uint32),
                ("vec_stride", ctypes.c_uint32), ("pos", ctypes.POINTER(ctypes.c_uint32)),
                ("mat", ctypes.POINTER(ctypes.c_uint8)), ("nmat", ctypes.c_uint32),
                ("grid_unit", ctypes.c_double), ("wave_num", ctypes.c_double),
                ("dmatrix", ctypes.POINTER(ctypes.c_double))]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint64), ("matvecs", ctypes.c_uint64), ("refinements", ctypes.c_uint64),
                ("rel_residual", ctypes.c_double), ("true_rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("status", ctypes.c_int32), ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("slots", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("cache_bytes", ctypes.c_size_t), ("solver_kind", ctypes.c_int), ("mode", ctypes.c_int),
                ("eps", ctypes.c_double), ("max_iter", ctypes.c_uint64), ("max_pack", ctypes.c_int),
                ("pack_grids", ctypes.c_int), ("long_solve_s", ctypes.c_double),
                ("long_solve_share", ctypes.c_double)]


class JobDesc(ctypes.Structure):
    _fields_ = [("geometry", OpDesc), ("nrhs", ctypes.c_int), ("cc_sqrt", ctypes.POINTER(ctypes.c_double)),
                ("einc", ctypes.POINTER(ctypes.c_double))]


class JobResult(ctypes.Structure):
    _fields_ = [("t_submit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_end", ctypes.c_double),
                ("status", ctypes.c_int32), ("packed_rhs", ctypes.c_int32), ("footprint", ctypes.c_size_t),
                ("cache_hit", ctypes.c_int32), ("iterations", ctypes.c_int32), ("setup_s", ctypes.c_double),
                ("tensor_s", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libdda_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdda_hip.so is missing: run `python dda_amd/build.py` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, sz, u32, i32, u64, dbl = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int, ctypes.c_uint64, ctypes.c_double
        pp = ctypes.POINTER(ctypes.c_void_p)
        dp = ctypes.POINTER(ctypes.c_double)
        L.dda_last_error.restype = ctypes.c_char_p
        L.dda_ctx_create.argtypes = [i32, pp]
        L.dda_ctx_destroy.argtypes = [vp]
        L.dda_ctx_synchronize.argtypes = [vp]
        L.dda_ctx_info.argtypes = [vp, ctypes.c_char_p, sz, ctypes.POINTER(sz), ctypes.POINTER(i32)]
        L.dda_malloc.argtypes = [vp, sz, pp]
        L.dda_free.argtypes = [vp, vp]
        L.dda_memcpy_h2d.argtypes = [vp, vp, vp, sz]
        L.dda_memcpy_d2h.argtypes = [vp, vp, vp, sz]
        L.dda_memcpy_d2d.argtypes = [vp, vp, vp, sz]
        L.dda_fft_fit.argtypes = [u32, u32, i32]
        L.dda_fft_fit_static.argtypes = [u32]
        L.dda_fft_fit_static.restype = u32
        L.dda_fft_fit.restype = u32
        L.dda_grid_plan.argtypes = [ctypes.POINTER(u32), i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.dda_op_create.argtypes = [vp, ctypes.POINTER(OpDesc), i32, pp]
        L.dda_op_destroy.argtypes = [vp]
        L.dda_op_info.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(sz)]
        L.dda_op_set_cc_sqrt.argtypes = [vp, dp]
        L.dda_op_set_cc_sqrt_batch.argtypes = [vp, dp, i32]
        L.dda_op_apply.argtypes = [vp, vp, vp, i32, vp]
        L.dda_op_apply_batch.argtypes = [vp, vp, vp, i32, i32, vp]
        L.dda_op_tune.argtypes = [vp, i32, ctypes.c_char_p, sz]
        L.dda_op_apply_host.argtypes = [vp, dp, dp, i32]
        L.dda_op_get_dmatrix.argtypes = [vp, dp]
        L.dda_op_get_variants.argtypes = [vp, ctypes.POINTER(i32)]
        L.dda_op_set_variants.argtypes = [vp, ctypes.POINTER(i32)]
        L.dda_op_profile.argtypes = [vp, i32]
        L.dda_op_profile_read.argtypes = [vp, dp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.dda_solver_create.argtypes = [vp, i32, u64, u64, pp]
        L.dda_solver_create_ex.argtypes = [vp, i32, u64, u64, i32, pp]
        L.dda_solver_set_inner_floor.argtypes = [vp, dbl]
        L.dda_solver_set_inner_plateau.argtypes = [vp, u64, dbl]
        L.dda_solver_set_gpu_share.argtypes = [vp, dbl, dbl]
        L.dda_solver_profile_read.argtypes = [vp, dp, ctypes.POINTER(u64), ctypes.POINTER(u64)]
        L.dda_solver_destroy.argtypes = [vp]
        L.dda_solver_profiling_run.argtypes = [vp, dp, dp, dp]
        L.dda_solver_solve.argtypes = [vp, dp, dbl, dp, dp, ctypes.POINTER(SolveInfo)]
        L.dda_solver_solve_batch.argtypes = [vp, dp, i32, dbl, dp, dp, ctypes.POINTER(SolveInfo)]
        L.dda_solver_history.argtypes = [vp, dp, u64, ctypes.POINTER(u64)]
        L.dda_dhat_cache_create.argtypes = [vp, sz, pp]
        L.dda_dhat_cache_destroy.argtypes = [vp]
        L.dda_dhat_cache_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(sz), ctypes.POINTER(i32)]
        L.dda_op_create_cached.argtypes = [vp, ctypes.POINTER(OpDesc), i32, vp, pp]
        L.dda_server_create.argtypes = [ctypes.POINTER(ServerConfig), pp]
        L.dda_server_submit.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(u64)]
        L.dda_server_wait.argtypes = [vp, u64, dp, ctypes.POINTER(SolveInfo), ctypes.POINTER(JobResult)]
        L.dda_server_stats.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(sz)]
        L.dda_server_destroy.argtypes = [vp]
        L.dda_solver_solve_batch_dev.argtypes = [vp, vp, i32, dbl, vp, vp, ctypes.POINTER(SolveInfo)]
        L.dda_couple_constants.argtypes = [i32, u32, dp, dbl, dbl, dp, dp, dp, dp]
        L.dda_couple_constants_diag.argtypes = [i32, u32, dp, dbl, dbl, dp, dp, dp, dp]
        L.dda_incident_plane.argtypes = [vp, i32, dp, dp, vp]
        L.dda_cross_sections.argtypes = [vp, vp, vp, i32, dp, dp, dp]
        L.dda_cross_sections_noalloc.argtypes = [vp, vp, vp, i32, dp, dp, dp]
        L.dda_far_field.argtypes = [vp, vp, i32, ctypes.POINTER(ctypes.c_int), dp, i32, dp]
        L.dda_jones.argtypes = [vp, vp, i32, dp, dp, dp, u32, dbl, dbl, u32, dbl, dbl, dp]
        L.dda_op_create_pack.argtypes = [vp, ctypes.POINTER(OpDesc), i32, i32, vp, pp]
        L.dda_op_pack_info.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(u32), ctypes.POINTER(u32)]
        L.dda_op_pack_map.argtypes = [vp, i32, ctypes.POINTER(u32)]
        L.dda_op_set_rhs_members.argtypes = [vp, ctypes.POINTER(i32), i32]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("dda_last_error", "dda_fft_fit", "dda_fft_fit_static"):
                f.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(status, what=""):
    if status != OK:
        msg = lib().dda_last_error().decode(errors="replace")
        cls = SolverError if status in (BREAKDOWN, MAX_ITER, STAGNATION) else DDAError
        raise cls(status, (what + ": " if what else "") + msg)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fft_fit(x, override=0, support_non_pot=True):
    """DDA/DDAParams.cpp:45-90 (x = 2 * box extent)."""
    return int(lib().dda_fft_fit(x, override, int(support_non_pot)))


def grid_plan(grid, precision):
    """(static_x, static_z): whether an operator on `grid` runs the X passes and
    the Z pass on compile-time planned kernels (host-only query)."""
    g = (ctypes.c_uint32 * 3)(*[int(v) for v in grid])
    sx, sz = ctypes.c_int(), ctypes.c_int()
    _check(lib().dda_grid_plan(g, int(precision), ctypes.byref(sx), ctypes.byref(sz)), "dda_grid_plan")
    return bool(sx.value), bool(sz.value)


_live_contexts = weakref.WeakSet()


class Context:
    """One GPU and one in-order stream (DDA/DDAMain.cpp:556-575).

    Objects created on a context (vectors, operators, solvers) register with it;
    closing the context releases them first, so no handle outlives it."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().dda_ctx_create(device, ctypes.byref(h)), "dda_ctx_create")
        self.handle = h
        self._children = []
        _live_contexts.add(self)

    def _adopt(self, obj):
        if len(self._children) > 256:
            self._children = [r for r in self._children if r() is not None]
        self._children.append(weakref.ref(obj))

    def info(self):
        name = ctypes.create_string_buffer(64)
        hbm = ctypes.c_size_t()
        cus = ctypes.c_int()
        _check(lib().dda_ctx_info(self.handle, name, 64, ctypes.byref(hbm), ctypes.byref(cus)))
        return {"arch": name.value.decode(), "hbm_bytes": hbm.value, "cu_count": cus.value}

    def synchronize(self):
        _check(lib().dda_ctx_synchronize(self.handle))

    def close(self):
        if self.handle:
            # solvers before operators before plain buffers
            live = [r() for r in reversed(self._children)]
            for obj in live:
                if obj is not None:
                    obj.close()
            self._children = []
            lib().dda_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@atexit.register
def _close_all_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


class DeviceVector:
    """Plain device buffer of 3*vec_stride complex values of one precision."""

    def __init__(self, ctx, n, precision):
        self.ctx, self.n, self.precision = ctx, int(n), precision
        self.dtype = np.complex128 if precision == 64 else np.complex64
        self.nbytes = self.n * np.dtype(self.dtype).itemsize
        p = ctypes.c_void_p()
        _check(lib().dda_malloc(ctx.handle, self.nbytes, ctypes.byref(p)), "dda_malloc")
        self.ptr = p
        ctx._adopt(self)

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.size == self.n
        _check(lib().dda_memcpy_h2d(self.ctx.handle, self.ptr, h.ctypes.data, self.nbytes))

    def download(self):
        h = np.empty(self.n, self.dtype)
        _check(lib().dda_memcpy_d2h(self.ctx.handle, h.ctypes.data, self.ptr, self.nbytes))
        return h

    def copy_from(self, other):
        """Device-to-device copy of another vector of the same size."""
        assert other.nbytes == self.nbytes
        _check(lib().dda_memcpy_d2d(self.ctx.handle, self.ptr, other.ptr, self.nbytes))

    def close(self):
        if self.ptr and self.ctx.handle:
            lib().dda_free(self.ctx.handle, self.ptr)
        self.ptr = None

    free = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuMatVec:
    """The FFT-convolution operator A = I - sqrt(a) G sqrt(a) on the device.

    GpuMatVec::create (DDA/GpuMatVec.hpp:85-86). `pos` is [nv, 3] grid
    coordinates, `mat` [nv] material indices, `grid` 0 entries use fftFit's
    smooth rule. `dmatrix` optionally supplies the reference-layout D matrix
    (complex [Gx, Gz, Gy, 6], 1/m^3); otherwise it is generated on the GPU.
    """

    def __init__(self, ctx, pos, mat, box, grid_unit, wave_num, precision=64, grid=(0, 0, 0), nmat=None,
                 vec_stride=0, dmatrix=None, cache=None):
        self.ctx = ctx
        self.pos = np.ascontiguousarray(pos, np.uint32).reshape(-1, 3)
        self.nv = self.pos.shape[0]
        self.mat = np.ascontiguousarray(mat if mat is not None else np.zeros(self.nv), np.uint8)
        self.nmat = int(nmat if nmat is not None else (int(self.mat.max()) + 1 if self.nv else 1))
        self.precision = precision
        d = OpDesc()
        d.box[:] = [int(b) for b in box]
        d.grid[:] = [int(g) for g in grid]
        d.nv = self.nv
        d.vec_stride = int(vec_stride)
        d.pos = self.pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        d.mat = self.mat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        d.nmat = self.nmat
        d.grid_unit = float(grid_unit)
        d.wave_num = float(wave_num)
        self._dm = None
        if dmatrix is not None:
            self._dm = np.ascontiguousarray(dmatrix, np.complex128)
            d.dmatrix = _dp(self._dm.view(np.float64))
        h = ctypes.c_void_p()
        if cache is not None:
            _check(lib().dda_op_create_cached(ctx.handle, ctypes.byref(d), precision, cache.handle, ctypes.byref(h)),
                   "dda_op_create_cached")
        else:
            _check(lib().dda_op_create(ctx.handle, ctypes.byref(d), precision, ctypes.byref(h)), "dda_op_create")
        self.handle = h
        ctx._adopt(self)
        self._dm = None
        g = (ctypes.c_uint32 * 3)()
        vs = ctypes.c_uint32()
        nb = ctypes.c_size_t()
        _check(lib().dda_op_info(h, g, ctypes.byref(vs), ctypes.byref(nb)))
        self.grid = tuple(int(v) for v in g)
        self.vec_stride = int(vs.value)
        self.device_bytes = int(nb.value)
        self.box = tuple(int(b) for b in box)
        self.grid_unit, self.wave_num = float(grid_unit), float(wave_num)

    @property
    def vec_size(self):
        return 3 * self.vec_stride

    def set_couple_constants(self, cc_sqrt):
        """GpuMatVec::setCoupleConstants: cc_sqrt [nmat, 3] complex (m^(3/2))."""
        cs = np.ascontiguousarray(np.asarray(cc_sqrt, np.complex128).reshape(self.nmat, 3))
        _check(lib().dda_op_set_cc_sqrt(self.handle, _dp(cs.view(np.float64))), "dda_op_set_cc_sqrt")

    def set_couple_constants_batch(self, cc_sqrt):
        """One set of couple constants per right-hand side: [nrhs, nmat, 3] complex."""
        cs = np.asarray(cc_sqrt, np.complex128)
        nrhs = cs.shape[0]
        cs = np.ascontiguousarray(cs.reshape(nrhs, self.nmat, 3))
        _check(lib().dda_op_set_cc_sqrt_batch(self.handle, _dp(cs.view(np.float64)), nrhs),
               "dda_op_set_cc_sqrt_batch")

    def apply(self, x, y, conj=False, stream=None):
        """Device vectors (DeviceVector or raw pointers); enqueued, not synchronised."""
        xp = x.ptr if isinstance(x, DeviceVector) else x
        yp = y.ptr if isinstance(y, DeviceVector) else y
        _check(lib().dda_op_apply(self.handle, xp, yp, int(conj), stream), "dda_op_apply")

    def apply_batch(self, x, y, nrhs, conj=False, stream=None):
        """Block matvec over nrhs vectors stored back to back ([nrhs][3][vec_stride])."""
        xp = x.ptr if isinstance(x, DeviceVector) else x
        yp = y.ptr if isinstance(y, DeviceVector) else y
        _check(lib().dda_op_apply_batch(self.handle, xp, yp, int(nrhs), int(conj), stream), "dda_op_apply_batch")

    def tune(self, nrhs=1):
        """Autotune the kernel variants for this shape (cached per process); nrhs >= 2 tunes
        the separate set used by blocks of right-hand sides. Returns the timing summary."""
        buf = ctypes.create_string_buffer(512)
        _check(lib().dda_op_tune(self.handle, int(nrhs), buf, 512), "dda_op_tune")
        return buf.value.decode()

    def variants(self, values=None):
        """Kernel variant per pass (Y fwd, Z, Y inv, X fwd, X inv) and the two-stream
        block split; with `values`, pins them first."""
        a = (ctypes.c_int32 * 6)()
        if values is not None:
            a[:] = [int(v) for v in values]
            _check(lib().dda_op_set_variants(self.handle, a), "dda_op_set_variants")
        _check(lib().dda_op_get_variants(self.handle, a), "dda_op_get_variants")
        return list(a)

    def dmatrix(self):
        """The device tensor in the reference layout [Gx, Gz, Gy, 6] (1/m^3)."""
        gx, gy, gz = self.grid
        out = np.zeros((gx, gz, gy, 6), np.complex128)
        _check(lib().dda_op_get_dmatrix(self.handle, _dp(out.view(np.float64))))
        return out

    def profile(self, enable=True):
        """Time each of the five kernels with HIP events on their own stream."""
        _check(lib().dda_op_profile(self.handle, int(enable)))

    def profile_read(self):
        """-> (ms per pass [5], launches per pass [5], skipped no-op launches)."""
        ms = np.zeros(5)
        cnt = (ctypes.c_uint64 * 5)()
        sk = ctypes.c_uint64()
        _check(lib().dda_op_profile_read(self.handle, _dp(ms), cnt, ctypes.byref(sk)))
        return ms, np.array(list(cnt), np.int64), int(sk.value)

    def close(self):
        if getattr(self, "handle", None):
            if self.ctx.handle:
                lib().dda_op_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PackMatVec(GpuMatVec):
    """Several particles on one grid in one operator (dda_op_create_pack).

    `members` is a list of (pos, mat, box, wave_num, nmat) sharing `grid_unit`
    and `grid`. The operator works on the union of the members' positions;
    `scatter` / `gather` move a member's [3, vec_stride_m] vector in and out of
    the union layout, and `members_of` assigns each right-hand side of a batch
    to a member. Couple constants are per right-hand side over the union's
    materials (the last one is void and forced to 0): `pad_cc` widens a
    member's [nmat_m, 3] set.
    """

    def __init__(self, ctx, members, grid_unit, precision=64, grid=(0, 0, 0), cache=None):
        self.ctx = ctx
        self.precision = precision
        self._keep = []
        descs = (OpDesc * len(members))()
        for d, (pos, mat, box, wave_num, nmat) in zip(descs, members):
            pos = np.ascontiguousarray(pos, np.uint32).reshape(-1, 3)
            mat = np.ascontiguousarray(mat if mat is not None else np.zeros(len(pos)), np.uint8)
            self._keep += [pos, mat]
            d.box[:] = [int(b) for b in box]
            d.grid[:] = [int(g) for g in grid]
            d.nv = len(pos)
            d.pos = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            d.mat = mat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
            d.nmat = int(nmat if nmat is not None else int(mat.max()) + 1)
            d.grid_unit = float(grid_unit)
            d.wave_num = float(wave_num)
        h = ctypes.c_void_p()
        _check(lib().dda_op_create_pack(ctx.handle, descs, len(members), precision,
                                        cache.handle if cache is not None else None, ctypes.byref(h)),
               "dda_op_create_pack")
        self.handle = h
        ctx._adopt(self)
        nm, nvu, nmu = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().dda_op_pack_info(h, ctypes.byref(nm), ctypes.byref(nvu), ctypes.byref(nmu)))
        self.nmember, self.nv, self.nmat = int(nm.value), int(nvu.value), int(nmu.value)
        self.member_nv = [int(d.nv) for d in descs]
        self.member_nmat = [int(d.nmat) for d in descs]
        self.maps = []
        for m in range(self.nmember):
            idx = np.zeros(self.member_nv[m], np.uint32)
            _check(lib().dda_op_pack_map(h, m, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))))
            self.maps.append(idx.astype(np.int64))
        self._keep = []
        g = (ctypes.c_uint32 * 3)()
        vs = ctypes.c_uint32()
        nb = ctypes.c_size_t()
        _check(lib().dda_op_info(h, g, ctypes.byref(vs), ctypes.byref(nb)))
        self.grid = tuple(int(v) for v in g)
        self.vec_stride = int(vs.value)
        self.device_bytes = int(nb.value)
        self.box = tuple(max(int(d.box[a]) for d in descs) for a in range(3))
        self.grid_unit = float(grid_unit)
        self.wave_nums = [float(d.wave_num) for d in descs]
        self.wave_num = self.wave_nums[0]

    def members_of(self, member):
        """Member of each right-hand side (a sequence of member indices)."""
        mm = np.ascontiguousarray(member, np.int32)
        _check(lib().dda_op_set_rhs_members(self.handle, mm.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(mm)),
               "dda_op_set_rhs_members")

    def pad_cc(self, member, cc_sqrt):
        """A member's [nmat_m, 3] couple constants widened to the union's materials."""
        out = np.zeros((self.nmat, 3), np.complex128)
        cs = np.asarray(cc_sqrt, np.complex128).reshape(-1, 3)
        out[:len(cs)] = cs
        return out

    def scatter(self, member, v):
        """A member vector [3, >= nv_m] -> union layout [3, vec_stride] (zero elsewhere)."""
        v = np.asarray(v).reshape(3, -1)
        out = np.zeros((3, self.vec_stride), np.result_type(v.dtype, np.complex64))
        out[:, self.maps[member]] = v[:, :self.member_nv[member]]
        return out

    def gather(self, member, u, stride=None):
        """Union vector [3, vec_stride] -> the member's [3, stride] (padding zero)."""
        u = np.asarray(u).reshape(3, -1)
        n = self.member_nv[member]
        out = np.zeros((3, stride or n), u.dtype)
        out[:, :n] = u[:, self.maps[member]]
        return out


class MatVecGpu:
    """MatVec<T> on host vectors (DDA/MatVecGpu.cpp:51-68): apply(arg, result, conj)."""

    def __init__(self, gpu_matvec):
        self.op = gpu_matvec

    def apply(self, arg, result=None, conj=False):
        x = np.ascontiguousarray(arg, np.complex128)
        assert x.size == self.op.vec_size
        y = np.empty_like(x) if result is None else result
        _check(lib().dda_op_apply_host(self.op.handle, _dp(x.view(np.float64)), _dp(y.view(np.float64)), int(conj)),
               "dda_op_apply_host")
        return y


class IterativeSolver:
    """IterativeSolverBase<T> over a GpuMatVec: getPolVec(einc, eps, start).

    mode "fp64" (default) runs everything in FP64 on an FP64 operator;
    "fp32" is the reference's float path (GpuQmrCs<float> over GpuMatVec<float>):
    the operator may be precision 32 or 64 (an FP32 twin is built); "mixed"
    (or mixed=True) runs the Krylov iteration in FP32 inside FP64 iterative
    refinement (no counterpart in the reference); `inner_floor` is the
    tolerance floor of the FP32 inner solves, `inner_plateau` = (window,
    ratio) their plateau stop."""

    kind = QMR_CS

    def __init__(self, gpu_matvec, max_iter=0, max_res_increase=0, mixed=False, inner_floor=None, mode=None,
                 inner_plateau=None):
        self.op = gpu_matvec
        mode = mode if mode is not None else ("mixed" if mixed else "fp64")
        self.mode = mode
        self.mixed = mode == "mixed"
        h = ctypes.c_void_p()
        _check(lib().dda_solver_create_ex(gpu_matvec.handle, self.kind, int(max_iter), int(max_res_increase),
                                          MODES[mode], ctypes.byref(h)), "dda_solver_create")
        if inner_floor is not None:
            _check(lib().dda_solver_set_inner_floor(h, float(inner_floor)))
        if inner_plateau is not None:  # (window, ratio)
            _check(lib().dda_solver_set_inner_plateau(h, int(inner_plateau[0]), float(inner_plateau[1])))
        self.handle = h
        self.info = None
        gpu_matvec.ctx._adopt(self)

    def set_gpu_share(self, after_seconds, share):
        """dda_solver_set_gpu_share: past after_seconds of a solve, leave the
        GPU to other streams for (1 / share - 1) of the time (share 1 = off)."""
        _check(lib().dda_solver_set_gpu_share(self.handle, float(after_seconds), float(share)),
               "dda_solver_set_gpu_share")

    def set_couple_constants(self, cc_sqrt):
        self.op.set_couple_constants(cc_sqrt)

    def get_pol_vec(self, einc, eps, start=None):
        e = np.ascontiguousarray(einc, np.complex128)
        assert e.size == self.op.vec_size
        out = np.zeros_like(e)
        st = None if start is None else np.ascontiguousarray(start, np.complex128)
        info = SolveInfo()
        status = lib().dda_solver_solve(self.handle, _dp(e.view(np.float64)), float(eps),
                                        None if st is None else _dp(st.view(np.float64)),
                                        _dp(out.view(np.float64)), ctypes.byref(info))
        self.info = info.as_dict()
        _check(status, "getPolVec")
        return out

    getPolVec = get_pol_vec

    def solve_device(self, einc_dev, pol_dev, nrhs, eps, start_dev=None):
        """Device-resident batched solve: einc_dev / pol_dev (/ start_dev) are
        DeviceVectors (FP64) of nrhs * vec_size entries. Returns the infos."""
        infos = (SolveInfo * nrhs)()
        status = lib().dda_solver_solve_batch_dev(self.handle, einc_dev.ptr, int(nrhs), float(eps),
                                                  start_dev.ptr if start_dev is not None else None, pol_dev.ptr,
                                                  infos)
        self.infos = [i.as_dict() for i in infos]
        self.info = self.infos[0]
        if status not in (OK, BREAKDOWN, MAX_ITER, STAGNATION, INVALID_ARG) or (
                status == INVALID_ARG and not any(i["status"] == INVALID_ARG for i in self.infos)):
            _check(status, "solve_device")
        return self.infos

    def get_pol_vecs(self, einc, eps, start=None):
        """Solves a batch of incident fields [nrhs, 3*vec_stride] in lockstep.

        Returns (pol [nrhs, 3*vec_stride], infos: list of dicts with a per-RHS
        status); raises only for errors of the call itself."""
        e = np.ascontiguousarray(einc, np.complex128)
        nrhs = e.shape[0]
        assert e.shape[1] == self.op.vec_size
        out = np.zeros_like(e)
        st = None if start is None else np.ascontiguousarray(start, np.complex128)
        infos = (SolveInfo * nrhs)()
        status = lib().dda_solver_solve_batch(self.handle, _dp(e.view(np.float64)), nrhs, float(eps),
                                              None if st is None else _dp(st.view(np.float64)),
                                              _dp(out.view(np.float64)), infos)
        self.infos = [i.as_dict() for i in infos]
        self.info = self.infos[0]
        if status not in (OK, BREAKDOWN, MAX_ITER, STAGNATION, INVALID_ARG) or (
                status == INVALID_ARG and not any(i["status"] == INVALID_ARG for i in self.infos)):
            _check(status, "get_pol_vecs")
        return out, self.infos

    def profile_read(self):
        """Kernel times of the operator the iteration runs on (FP32 twin when mixed)."""
        ms = np.zeros(5)
        cnt = (ctypes.c_uint64 * 5)()
        sk = ctypes.c_uint64()
        _check(lib().dda_solver_profile_read(self.handle, _dp(ms), cnt, ctypes.byref(sk)))
        return ms, np.array(list(cnt), np.int64), int(sk.value)

    def profiling_run(self):
        """IterativeSolverBase::profilingRun (IterativeSolverBase.cpp:105-136):
        five iterations timed one by one -> (microseconds in run order, median,
        spread)."""
        us = np.zeros(5)
        med, spread = ctypes.c_double(), ctypes.c_double()
        _check(lib().dda_solver_profiling_run(self.handle, _dp(us), ctypes.byref(med), ctypes.byref(spread)),
               "profilingRun")
        return us, med.value, spread.value

    def history(self):
        n = int(self.info["iterations"]) + 1 if self.info else 1
        out = np.zeros(n)
        w = ctypes.c_uint64()
        _check(lib().dda_solver_history(self.handle, _dp(out), n, ctypes.byref(w)))
        return out[: w.value]

    def close(self):
        if getattr(self, "handle", None):
            if self.op.handle and self.op.ctx.handle:
                lib().dda_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuQmrCs(IterativeSolver):
    kind = QMR_CS


class GpuBicgStab(IterativeSolver):
    kind = BICGSTAB


class GpuCgnr(IterativeSolver):
    kind = CGNR


class GpuBicgCs(IterativeSolver):
    kind = BICG_CS


SOLVER_CLASSES = {"qmr": GpuQmrCs, "bicgstab": GpuBicgStab, "cgnr": GpuCgnr, "bicg": GpuBicgCs}


class TensorCache:
    """Device cache of FFT'd interaction tensors keyed by (grid, kd, precision)."""

    def __init__(self, ctx, max_bytes=1 << 34):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().dda_dhat_cache_create(ctx.handle, int(max_bytes), ctypes.byref(h)), "dda_dhat_cache_create")
        self.handle = h

    def stats(self):
        c = (ctypes.c_uint64 * 3)()
        b = ctypes.c_size_t()
        e = ctypes.c_int()
        _check(lib().dda_dhat_cache_stats(self.handle, c, ctypes.byref(b), ctypes.byref(e)))
        return {"hits": c[0], "misses": c[1], "evictions": c[2], "bytes": b.value, "entries": e.value}

    def close(self):
        if getattr(self, "handle", None):
            _check(lib().dda_dhat_cache_destroy(self.handle), "dda_dhat_cache_destroy")
            self.handle = None


class Server:
    """Scattering-job server (dda_server_*): queue, HBM admission, `slots`
    concurrent block solves on their own streams, tensor cache, packing of
    jobs of the same particle (with pack_grids, of any particles on one grid).
    long_solve = (seconds, share): every job's solver yields the GPU to the
    other slots once it has run that long (dda_solver_set_gpu_share)."""

    def __init__(self, device=0, slots=4, solver="bicgstab", mixed=True, eps=1e-6, max_iter=0, max_pack=16,
                 hbm_budget=0, cache_bytes=0, pack_grids=False, long_solve=None):
        c = ServerConfig()
        c.pack_grids = int(bool(pack_grids))
        if long_solve is not None:
            c.long_solve_s, c.long_solve_share = float(long_solve[0]), float(long_solve[1])
        c.device, c.slots, c.hbm_budget, c.cache_bytes = device, slots, int(hbm_budget), int(cache_bytes)
        c.solver_kind = SOLVER_KINDS[solver]
        c.mode = MODE_MIXED if mixed else MODE_FP64
        c.eps, c.max_iter, c.max_pack = float(eps), int(max_iter), int(max_pack)
        h = ctypes.c_void_p()
        _check(lib().dda_server_create(ctypes.byref(c), ctypes.byref(h)), "dda_server_create")
        self.handle = h
        self._shapes = {}

    def submit(self, pos, mat, box, grid_unit, wave_num, einc, cc_sqrt, nmat=None, grid=(0, 0, 0)):
        """einc [nrhs, 3*vec_stride] complex, cc_sqrt [nrhs, nmat, 3] complex -> job id."""
        pos = np.ascontiguousarray(pos, np.uint32).reshape(-1, 3)
        nv = pos.shape[0]
        mat = np.ascontiguousarray(mat if mat is not None else np.zeros(nv), np.uint8)
        nmat = int(nmat if nmat is not None else int(mat.max()) + 1)
        e = np.ascontiguousarray(einc, np.complex128)
        e = e.reshape(-1, e.shape[-1])
        cs = np.ascontiguousarray(np.asarray(cc_sqrt, np.complex128).reshape(e.shape[0], nmat, 3))
        j = JobDesc()
        g = j.geometry
        g.box[:] = [int(b) for b in box]
        g.grid[:] = [int(x) for x in grid]
        g.nv = nv
        g.vec_stride = e.shape[1] // 3
        g.pos = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        g.mat = mat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        g.nmat = nmat
        g.grid_unit, g.wave_num = float(grid_unit), float(wave_num)
        j.nrhs = e.shape[0]
        j.cc_sqrt = _dp(cs.view(np.float64))
        j.einc = _dp(e.view(np.float64))
        jid = ctypes.c_uint64()
        _check(lib().dda_server_submit(self.handle, ctypes.byref(j), ctypes.byref(jid)), "dda_server_submit")
        self._shapes[jid.value] = e.shape
        return jid.value

    def wait(self, job_id):
        """-> (pol [nrhs, 3*vec_stride], infos, result dict); status in result["status"]."""
        shape = self._shapes.pop(job_id)
        pol = np.zeros(shape, np.complex128)
        infos = (SolveInfo * shape[0])()
        res = JobResult()
        st = lib().dda_server_wait(self.handle, job_id, _dp(pol.view(np.float64)), infos, ctypes.byref(res))
        if st not in (OK, BREAKDOWN, MAX_ITER, STAGNATION, INVALID_ARG):
            _check(st, "dda_server_wait")
        return pol, [i.as_dict() for i in infos], res.as_dict()

    def stats(self):
        c = (ctypes.c_uint64 * 5)()
        pk = ctypes.c_size_t()
        _check(lib().dda_server_stats(self.handle, c, ctypes.byref(pk)))
        return {"submitted": c[0], "completed": c[1], "cache_hits": c[2], "cache_misses": c[3],
                "cache_evictions": c[4], "peak_bytes": pk.value}

    def close(self):
        if getattr(self, "handle", None):
            lib().dda_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def couple_constants(materials, grid_unit, kd, prop, pol, kind="ldr"):
    """dda_couple_constants -> (cc, cc_sqrt), each [nmat, 3] complex. A material
    given as three values (the diagonal of an anisotropic index) goes through
    dda_couple_constants_diag."""
    aniso = any(np.ndim(v) > 0 for v in materials)
    if aniso:
        m = np.ascontiguousarray([np.broadcast_to(np.asarray(v, np.complex128), (3,)) for v in materials])
        n = m.shape[0]
        fn, name = lib().dda_couple_constants_diag, "dda_couple_constants_diag"
    else:
        m = np.ascontiguousarray(np.asarray(materials, np.complex128).reshape(-1))
        n = m.size
        fn, name = lib().dda_couple_constants, "dda_couple_constants"
    cc = np.zeros((n, 3), np.complex128)
    cs = np.zeros((n, 3), np.complex128)
    pr = np.ascontiguousarray(prop, np.float64)
    pl = np.ascontiguousarray(pol, np.float64)
    _check(fn(POL_KINDS[kind], n, _dp(m.view(np.float64)), float(grid_unit), float(kd), _dp(pr), _dp(pl),
              _dp(cc.view(np.float64)), _dp(cs.view(np.float64))), name)
    return cc, cs


def incident_plane(op, props, pols, out=None):
    """Plane-wave incident fields on the device for each (prop, pol) row -> DeviceVector [nrhs * vec_size]."""
    pr = np.ascontiguousarray(np.asarray(props, np.float64).reshape(-1, 3))
    pl = np.ascontiguousarray(np.asarray(pols, np.float64).reshape(-1, 3))
    nrhs = pr.shape[0]
    if out is None:
        out = DeviceVector(op.ctx, nrhs * op.vec_size, 64)
    _check(lib().dda_incident_plane(op.handle, nrhs, _dp(pr), _dp(pl), out.ptr), "dda_incident_plane")
    return out


def cross_sections(op, pol_dev, einc_dev, nrhs, cc=None):
    """(cext [nrhs], cabs [nrhs] or None) from device vectors; cc [nrhs, nmat, 3] alpha."""
    cext = np.zeros(nrhs)
    cabs = np.zeros(nrhs) if cc is not None else None
    c = None if cc is None else np.ascontiguousarray(np.asarray(cc, np.complex128).reshape(nrhs, -1, 3))
    _check(lib().dda_cross_sections(op.handle, pol_dev.ptr, einc_dev.ptr, int(nrhs),
                                    None if c is None else _dp(c.view(np.float64)), _dp(cext),
                                    None if cabs is None else _dp(cabs)), "dda_cross_sections")
    return cext, cabs


def cross_sections_noalloc(op, pol_dev, einc_dev, nrhs, cc=None):
    """cross_sections through dda_cross_sections_noalloc (no device allocation:
    the library's fixed per-GPU scratch array; overlapping calls on one GPU race)."""
    cext = np.zeros(nrhs)
    cabs = np.zeros(nrhs) if cc is not None else None
    c = None if cc is None else np.ascontiguousarray(np.asarray(cc, np.complex128).reshape(nrhs, -1, 3))
    _check(lib().dda_cross_sections_noalloc(op.handle, pol_dev.ptr, einc_dev.ptr, int(nrhs),
                                            None if c is None else _dp(c.view(np.float64)), _dp(cext),
                                            None if cabs is None else _dp(cabs)), "dda_cross_sections_noalloc")
    return cext, cabs


def far_field(op, pol_dev, nrhs, dirs, rhs=None):
    """Scattered far fields for directions dirs [nq, 3] (of right-hand side rhs[q]) -> [nq, 3] complex."""
    d = np.ascontiguousarray(np.asarray(dirs, np.float64).reshape(-1, 3))
    nq = d.shape[0]
    r = None if rhs is None else np.ascontiguousarray(rhs, np.int32)
    out = np.zeros((nq, 3), np.complex128)
    _check(lib().dda_far_field(op.handle, pol_dev.ptr, int(nrhs),
                               None if r is None else r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(d), nq,
                               _dp(out.view(np.float64))), "dda_far_field")
    return out


def jones(op, pol_dev, symmetric, prop, pol1, pol2, ntheta, theta0, dtheta, nphi, phi0, dphi):
    """Jones matrices [ntheta * nphi, 2, 2] complex from the device solutions of both polarisations."""
    out = np.zeros((ntheta * nphi, 2, 2), np.complex128)
    a = [np.ascontiguousarray(v, np.float64) for v in (prop, pol1, pol2)]
    _check(lib().dda_jones(op.handle, pol_dev.ptr, int(symmetric), _dp(a[0]), _dp(a[1]), _dp(a[2]), int(ntheta),
                           float(theta0), float(dtheta), int(nphi), float(phi0), float(dphi),
                           _dp(out.view(np.float64))), "dda_jones")
    return out


def solve_particle(solver, params, prop, eps, kind="ldr", symmetric=None, pol_dev=None):
    """Both plane-wave polarisations of one particle, as the reference's
    createResOutput solves them (DDA/DDAMain.cpp:380-410, 452-470): for a
    rotation-symmetric target (problem.rotation_symmetric; `symmetric`
    overrides the rule) polarisation 1 alone is solved, polarisation 2's cross
    sections are polarisation 1's and its solution is polarisation 1's turned
    a quarter turn about z (jones(..., symmetric=1) applies that); otherwise
    both are solved as one 2-RHS block. Incident fields are built on the
    device. Returns a dict: symmetric, nrhs, pol_dev (DeviceVector of nrhs
    solutions), infos, cext [2], cabs [2], prop, pol1, pol2."""
    from . import problem as P
    op = solver.op
    prop_n, pol1, pol2 = P.beam_pols(prop)
    sym = P.rotation_symmetric(params.geometry, prop_n) if symmetric is None else bool(symmetric)
    pols = [pol1] if sym else [pol1, pol2]
    ccs = [P.couple_constants(params, prop_n, p, kind) for p in pols]
    k = len(pols)
    op.set_couple_constants_batch(np.array([cs for _, cs in ccs]))
    e_dev = incident_plane(op, [prop_n] * k, pols)
    if pol_dev is None:
        pol_dev = DeviceVector(op.ctx, k * op.vec_size, 64)
    try:
        infos = solver.solve_device(e_dev, pol_dev, k, eps)
        cext, cabs = cross_sections(op, pol_dev, e_dev, k, np.array([cc for cc, _ in ccs]))
    finally:
        e_dev.close()
    if sym:
        cext, cabs = np.array([cext[0], cext[0]]), np.array([cabs[0], cabs[0]])
    return {"symmetric": sym, "nrhs": k, "pol_dev": pol_dev, "infos": infos, "cext": cext, "cabs": cabs,
            "prop": prop_n, "pol1": pol1, "pol2": pol2}
