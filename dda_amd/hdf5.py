"""HDF5 output of dipole polarisations and Jones far fields in the reference's
file layout: MATLAB 7.3 files (HDF5 with a 512-byte "MAT" user block) written
the way HDF5::matlabSerialize writes them (HDF5/Matlab.cpp, HDF5/Matlab.hpp):

  struct            group, attribute MATLAB_class = "struct", one link per
                    member in declaration order; empty (null) members skipped;
                    the file's top-level struct is written into the root group
  string            uint16 (UTF-16) dataset, dims (len, 1), MATLAB_class "char",
                    MATLAB_int_decode = 2
  scalar            dims (1); vector<T> dims (n); Vector3 / DiagMatrix3 dims (3);
                    Quaternion dims (4); vector<Vector3> / vector<DiagMatrix3>
                    dims (n, 3); multi_array (Fortran order) dims reversed
  long double       stored as double ("double"); complex as the compound
                    {real, imag} of the element type
  attribute string  fixed-length C string of the value's length, scalar space
  objects           no time stamps (H5Pset_obj_track_times(false))

Files (the reference's names and contents):
  DipPol-PolN.hdf5  DDAFieldFile (EMSim/DataFilesDDA.hpp:110-123) from
                    createDDAFieldFile (DDA/DataFilesDDAUtil.cpp:112-120),
                    written by DDAFieldFile::write (EMSim/DataFilesDDA.cpp:58-63)
  <Far>.hdf5        JonesFarFieldFile (EMSim/DataFiles.hpp:155-171) from
                    JonesCalculus::store (EMSim/JonesCalculus.hpp:82-93) with the
                    dipole geometry without positions (FarFieldCalc.cpp:173)
with Parameters (EMSim/DataFiles.hpp:67-94; DDA/DataFilesDDAUtil.cpp:31-50 and
DDA/DDAMain.cpp:362-368) and the geometry (DDA/DataFilesDDAUtil.cpp:52-85).

The HDF5 C library is loaded with ctypes (DDA_HDF5_LIB, else libhdf5 from the
conda prefix or the loader's path); these functions raise if it is absent.
"""
import ctypes
import ctypes.util
import os

import numpy as np

SPEED_OF_LIGHT = 299792458.0  # DDA/FPConst.hpp:40

_L = None
hid_t = ctypes.c_int64
hsize_t = ctypes.c_uint64
herr_t = ctypes.c_int

H5F_ACC_TRUNC = 0x0002
H5F_ACC_RDONLY = 0x0000
H5P_DEFAULT = 0
H5S_SCALAR = 0
H5T_COMPOUND = 6
H5S_ALL = 0


def _lib():
    global _L
    if _L is not None:
        return _L
    cands = [os.environ.get("DDA_HDF5_LIB"), "/opt/conda/lib/libhdf5.so", ctypes.util.find_library("hdf5")]
    err = None
    for c in cands:
        if not c:
            continue
        try:
            L = ctypes.CDLL(c)
            break
        except OSError as e:
            err = e
    else:
        raise RuntimeError("HDF5 output needs the HDF5 C library (libhdf5): %s" % (err or "not found"))
    if L.H5open() < 0:
        raise RuntimeError("H5open failed")
    fns = {
        "H5Fcreate": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t, hid_t]),
        "H5Fopen": (hid_t, [ctypes.c_char_p, ctypes.c_uint, hid_t]),
        "H5Fclose": (herr_t, [hid_t]),
        "H5Pcreate": (hid_t, [hid_t]),
        "H5Pclose": (herr_t, [hid_t]),
        "H5Pset_userblock": (herr_t, [hid_t, hsize_t]),
        "H5Pset_obj_track_times": (herr_t, [hid_t, ctypes.c_uint]),
        "H5Gcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t]),
        "H5Gopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Gclose": (herr_t, [hid_t]),
        "H5Screate": (hid_t, [ctypes.c_int]),
        "H5Screate_simple": (hid_t, [ctypes.c_int, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sget_simple_extent_ndims": (ctypes.c_int, [hid_t]),
        "H5Sget_simple_extent_dims": (ctypes.c_int, [hid_t, ctypes.POINTER(hsize_t), ctypes.POINTER(hsize_t)]),
        "H5Sclose": (herr_t, [hid_t]),
        "H5Tcopy": (hid_t, [hid_t]),
        "H5Tcreate": (hid_t, [ctypes.c_int, ctypes.c_size_t]),
        "H5Tinsert": (herr_t, [hid_t, ctypes.c_char_p, ctypes.c_size_t, hid_t]),
        "H5Tset_size": (herr_t, [hid_t, ctypes.c_size_t]),
        "H5Tclose": (herr_t, [hid_t]),
        "H5Dcreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        "H5Dopen2": (hid_t, [hid_t, ctypes.c_char_p, hid_t]),
        "H5Dwrite": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dread": (herr_t, [hid_t, hid_t, hid_t, hid_t, hid_t, ctypes.c_void_p]),
        "H5Dget_space": (hid_t, [hid_t]),
        "H5Dclose": (herr_t, [hid_t]),
        "H5Acreate2": (hid_t, [hid_t, ctypes.c_char_p, hid_t, hid_t, hid_t, hid_t]),
        "H5Awrite": (herr_t, [hid_t, hid_t, ctypes.c_void_p]),
        "H5Aclose": (herr_t, [hid_t]),
        "H5Lexists": (ctypes.c_int, [hid_t, ctypes.c_char_p, hid_t]),
    }
    for name, (res, args) in fns.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    g = lambda n: hid_t.in_dll(L, n).value
    L.T = {"double": g("H5T_NATIVE_DOUBLE_g"), "single": g("H5T_NATIVE_FLOAT_g"), "uint8": g("H5T_NATIVE_UCHAR_g"),
           "uint16": g("H5T_NATIVE_USHORT_g"), "uint32": g("H5T_NATIVE_UINT_g"), "int32": g("H5T_NATIVE_INT_g"),
           "c_s1": g("H5T_C_S1_g")}
    L.P = {"fcpl": g("H5P_CLS_FILE_CREATE_ID_g"), "dcpl": g("H5P_CLS_DATASET_CREATE_ID_g"),
           "gcpl": g("H5P_CLS_GROUP_CREATE_ID_g")}
    _L = L
    return L


def _chk(v, what):
    if v < 0:
        raise RuntimeError("HDF5 call failed: %s" % what)
    return v


def available():
    try:
        _lib()
        return True
    except RuntimeError:
        return False


class _Writer:
    """The serialisation rules of HDF5/Matlab.cpp over the C API."""

    def __init__(self, L, fid):
        self.L, self.fid = L, fid
        self.dcpl = _chk(L.H5Pcreate(L.P["dcpl"]), "H5Pcreate")
        _chk(L.H5Pset_obj_track_times(self.dcpl, 0), "H5Pset_obj_track_times")
        self.gcpl = _chk(L.H5Pcreate(L.P["gcpl"]), "H5Pcreate")
        _chk(L.H5Pset_obj_track_times(self.gcpl, 0), "H5Pset_obj_track_times")
        self.cplx = {}

    def close(self):
        for t in self.cplx.values():
            self.L.H5Tclose(t)
        self.L.H5Pclose(self.dcpl)
        self.L.H5Pclose(self.gcpl)

    def complex_type(self, cls):
        if cls not in self.cplx:
            base = self.L.T[cls]
            sz = 8 if cls == "double" else 4
            t = _chk(self.L.H5Tcreate(H5T_COMPOUND, 2 * sz), "H5Tcreate")
            _chk(self.L.H5Tinsert(t, b"real", 0, base), "H5Tinsert")
            _chk(self.L.H5Tinsert(t, b"imag", sz, base), "H5Tinsert")
            self.cplx[cls] = t
        return self.cplx[cls]

    def attr_string(self, obj, name, value):
        L = self.L
        b = value.encode()
        t = _chk(L.H5Tcopy(L.T["c_s1"]), "H5Tcopy")
        _chk(L.H5Tset_size(t, len(b)), "H5Tset_size")
        sp = _chk(L.H5Screate(H5S_SCALAR), "H5Screate")
        a = _chk(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        buf = ctypes.create_string_buffer(b, len(b))
        _chk(L.H5Awrite(a, t, buf), "H5Awrite")
        L.H5Aclose(a)
        L.H5Sclose(sp)
        L.H5Tclose(t)

    def attr_int32(self, obj, name, value):
        L = self.L
        sp = _chk(L.H5Screate(H5S_SCALAR), "H5Screate")
        a = _chk(L.H5Acreate2(obj, name.encode(), L.T["int32"], sp, H5P_DEFAULT, H5P_DEFAULT), "H5Acreate2")
        v = ctypes.c_int32(value)
        _chk(L.H5Awrite(a, L.T["int32"], ctypes.byref(v)), "H5Awrite")
        L.H5Aclose(a)
        L.H5Sclose(sp)

    def group(self, parent, name):
        g = _chk(self.L.H5Gcreate2(parent, name.encode(), H5P_DEFAULT, self.gcpl, H5P_DEFAULT), "H5Gcreate2")
        self.attr_string(g, "MATLAB_class", "struct")
        return g

    def dataset(self, parent, name, data, cls, dims):
        """data: a numpy array in the file's element order; cls: MATLAB class."""
        L = self.L
        arr = np.ascontiguousarray(data)
        if np.iscomplexobj(arr):
            arr = np.ascontiguousarray(arr, np.complex128 if cls == "double" else np.complex64)
            t = self.complex_type(cls)
        else:
            npt = {"double": np.float64, "single": np.float32, "uint8": np.uint8, "uint16": np.uint16,
                   "uint32": np.uint32}[cls]
            arr = np.ascontiguousarray(arr, npt)
            t = L.T[cls]
        d = (hsize_t * len(dims))(*dims)
        sp = _chk(L.H5Screate_simple(len(dims), d, None), "H5Screate_simple")
        ds = _chk(L.H5Dcreate2(parent, name.encode(), t, sp, H5P_DEFAULT, self.dcpl, H5P_DEFAULT), "H5Dcreate2")
        self.attr_string(ds, "MATLAB_class", cls)
        if arr.size:
            _chk(L.H5Dwrite(ds, t, sp, sp, H5P_DEFAULT, arr.ctypes.data), "H5Dwrite")
        return ds, sp

    def put(self, parent, name, data, cls, dims):
        ds, sp = self.dataset(parent, name, data, cls, dims)
        self.L.H5Dclose(ds)
        self.L.H5Sclose(sp)

    def struct(self, parent, name, members):
        g = self.group(parent, name)
        self.members(g, members)
        self.L.H5Gclose(g)

    def members(self, g, members):
        for mname, kind, value in members:
            if value is None:
                continue  # null members are not linked (linkIfNotNull)
            if kind == "string":
                self.char(g, mname, value)
            elif kind == "struct":
                self.struct(g, mname, value)
            else:
                cls, dims = kind
                self.put(g, mname, value, cls, dims)

    def char(self, parent, name, s):
        L = self.L
        u = np.frombuffer(s.encode("utf-16-le"), np.uint16).copy() if s else np.zeros(0, np.uint16)
        d = (hsize_t * 2)(u.size, 1)
        sp = _chk(L.H5Screate_simple(2, d, None), "H5Screate_simple")
        ds = _chk(L.H5Dcreate2(parent, name.encode(), L.T["uint16"], sp, H5P_DEFAULT, self.dcpl, H5P_DEFAULT),
                  "H5Dcreate2")
        self.attr_string(ds, "MATLAB_class", "char")
        self.attr_int32(ds, "MATLAB_int_decode", 2)
        if u.size:
            _chk(L.H5Dwrite(ds, L.T["uint16"], sp, sp, H5P_DEFAULT, u.ctypes.data), "H5Dwrite")
        L.H5Dclose(ds)
        L.H5Sclose(sp)


def _write_file(path, members):
    """createMatlabFile (HDF5/Matlab.cpp:220-254) and the top-level struct
    written into the root group (matlabSerialize(file, object))."""
    L = _lib()
    fcpl = _chk(L.H5Pcreate(L.P["fcpl"]), "H5Pcreate")
    _chk(L.H5Pset_userblock(fcpl, 512), "H5Pset_userblock")
    fid = _chk(L.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, fcpl, H5P_DEFAULT), "H5Fcreate")
    L.H5Pclose(fcpl)
    w = _Writer(L, fid)
    try:
        root = _chk(L.H5Gopen2(fid, b"/", H5P_DEFAULT), "H5Gopen2")
        w.members(root, members)
        L.H5Gclose(root)
    finally:
        w.close()
        _chk(L.H5Fclose(fid), "H5Fclose")
    ub = bytearray(512)
    ub[0:3] = b"MAT"
    ub[125], ub[126], ub[127] = 2, ord("I"), ord("M")
    with open(path, "r+b") as f:
        f.write(bytes(ub))


def _vec3(v):
    return np.asarray(v, np.float64).reshape(3)


def _parameters(params, prop, pols, kind="ldr", geometry_string="", cmdline="", beam="plane", gamma=0.001):
    """Parameters<DDAParameters> (EMSim/DataFiles.hpp:67-94) as createParametersDDA
    and DDAMain fill them (DDA/DataFilesDDAUtil.cpp:31-50, DDA/DDAMain.cpp:362-368)."""
    method = [("PolarizabilityType", "string", kind),
              ("Gamma", ("double", (1,)), np.array([gamma])),
              ("Frequency", ("double", (1,)), np.array([SPEED_OF_LIGHT / params.lambda_]))]
    p = np.asarray(pols, np.float64).reshape(-1, 3)
    return [("Method", "string", "DDA"),
            ("MethodParameters", "struct", method),
            ("GeometryString", "string", geometry_string),
            ("CmdLine", "string", cmdline),
            ("RefractiveIndexMedium", ("double", (3,)), np.ones(3)),
            ("PropagationVector", ("double", (3,)), _vec3(prop)),
            ("BeamString", "string", beam),
            ("Polarizations", ("double", (p.shape[0], 3)), p)]


def _geometry(geometry, full):
    """DDADipoleListGeometry (EMSim/DataFilesDDA.hpp:65-93) as
    createDDADipoleListGeometry fills it (DDA/DataFilesDDAUtil.cpp:52-85):
    positions and material indices only with `full`, no periodicity."""
    g = geometry
    mats = np.array([np.broadcast_to(np.asarray(m, np.complex128), (3,)) for m in g.materials])
    return [("Type", "string", "DDADipoleList"),
            ("GridSpacing", ("double", (3,)), np.full(3, g.grid_unit)),
            ("Size", ("uint32", (3,)), np.asarray(g.box, np.uint32)),
            ("GridOrigin", ("double", (3,)), np.asarray(g.origin, np.float64)),
            ("Orientation", ("double", (4,)), np.array([1.0, 0.0, 0.0, 0.0])),
            ("RefractiveIndices", ("double", (mats.shape[0], 3)), mats),
            ("DipolePositions", ("uint32", (g.nv, 3)), g.pos if full else None),
            ("DipoleMaterialIndices", ("uint8", (g.nv,)), g.mat if full else None),
            ("Periodicity", ("double", (0, 3)), np.zeros((0, 3)))]


def write_dippol(path, params, pvec, beam_polarization, prop, pols, field_name="DipolePolarization", **par):
    """DipPol-PolN.hdf5: a DDAFieldFile of the dipole polarisation vector pvec
    (SoA [3 * vec_stride], FP64 -> "double", FP32 -> "single") for incident
    polarisation `beam_polarization` (1 or 2); prop and pols (2 x 3) go into
    the Parameters. `par`: kind, geometry_string, cmdline, beam, gamma."""
    v = np.asarray(pvec)
    cls = "single" if v.dtype == np.complex64 else "double"
    data = params.from_soa(v.astype(np.complex128 if cls == "double" else np.complex64))
    field = [("FieldName", "string", field_name),
             ("BeamPolarization", ("uint32", (1,)), np.array([beam_polarization], np.uint32)),
             ("Data", (cls, (params.nv, 3)), data)]
    _write_file(path, [("Type", "string", "DDAField"),
                       ("Parameters", "struct", _parameters(params, prop, pols, **par)),
                       ("Geometry", "struct", _geometry(params.geometry, True)),
                       ("Field", "struct", field)])


def write_jones(path, params, theta, phi, jones, prop, pols, **par):
    """A Jones far-field file (JonesFarFieldFile): theta, phi [n] (rad) per
    angle, jones [n][2][2] complex (getJonesMatrix, JonesCalculus.cpp:74-86),
    one frequency; Data is the Fortran-order [2][2][n][1] array of
    computeJonesFarField (JonesCalculus.cpp:50-72)."""
    J = np.asarray(jones)
    n = J.shape[0]
    cls = "single" if J.dtype == np.complex64 else "double"
    data = np.ascontiguousarray(J.transpose(0, 2, 1)).reshape(1, n, 2, 2)  # HDF5 dims reverse Fortran's
    jff = [("Theta", ("double", (n,)), np.asarray(theta, np.float64)),
           ("Phi", ("double", (n,)), np.asarray(phi, np.float64)),
           ("Frequency", ("double", (1,)), np.array([SPEED_OF_LIGHT / params.lambda_])),
           ("Data", (cls, (1, n, 2, 2)), data)]
    _write_file(path, [("Type", "string", "JonesFarField"),
                       ("Parameters", "struct", _parameters(params, prop, pols, **par)),
                       ("Geometry", "struct", _geometry(params.geometry, False)),
                       ("JonesFarField", "struct", jff)])


def read_dataset(path, name, complex_=False, dtype=np.float64):
    """Reads dataset `name` ("/Field/Data", ...) -> numpy array in the file's
    dims (complex compound {real, imag} with complex_)."""
    L = _lib()
    fid = _chk(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen")
    try:
        ds = _chk(L.H5Dopen2(fid, name.encode(), H5P_DEFAULT), "H5Dopen2")
        sp = L.H5Dget_space(ds)
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (hsize_t * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        shape = tuple(int(dims[i]) for i in range(nd))
        if complex_:
            out = np.zeros(shape, np.complex128)
            w = _Writer(L, fid)
            t = w.complex_type("double")
            if out.size:
                _chk(L.H5Dread(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread")
            w.close()
        else:
            out = np.zeros(shape, dtype)
            mt = {np.float64: "double", np.uint32: "uint32", np.uint8: "uint8", np.uint16: "uint16"}[dtype]
            if out.size:
                _chk(L.H5Dread(ds, L.T[mt], H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread")
        L.H5Sclose(sp)
        L.H5Dclose(ds)
    finally:
        L.H5Fclose(fid)
    return out


def read_dippol(path, params):
    """The polarisation vector (SoA [3 * vec_stride]) of a DipPol file, as
    Load::loadDipPol reads it back for --load-dip-pol (DDA/Load.cpp:76-107)."""
    d = read_dataset(path, "/Field/Data", complex_=True)
    if d.shape != (params.nv, 3):
        raise ValueError("DipPol file holds %s values, the geometry %d dipoles" % (d.shape, params.nv))
    return params.to_soa(d)
