"""Problem set-up around the matvec: geometry, grid, polarisabilities, beam.

These are the inputs of the hot path, restated from the reference:
  geometry    DDA/Shapes.cpp:53-143, DDA/DipoleGeometry.cpp:75-107, DDA/Load.cpp:37-69
  DDAParams   DDA/DDAParams.cpp:93-214 (grid via fftFit, vec_stride, k, kd)
  couple      DDA/PolarizabilityDescription.cpp:65-110, DDA/FPConst.hpp:48-50,
  constants   DDA/DDAParams.cpp:315-331
  beam        DDA/Beam.cpp:34-51 (polarisations), :110-118 (plane wave)
  observables DDA/Beam.cpp:58-68 (Cext), DDA/AbsCross.cpp:27-48 (Cabs),
              DDA/CpuFieldCalculator.cpp:48-98 (far field)
The ellipsoid is not a reference shape: it applies the sphere's cell-centre
rule (Shapes.cpp:59-65) per axis.
"""
import math

import numpy as np

from . import fft_fit

LDR_B1, LDR_B2, LDR_B3 = 1.8915316, -0.1648469, 1.7700004  # FPConst.hpp:48-50 (doi:10.1086/172396)


class DipoleGeometry:
    """Dipoles sorted by (z, y, x) with material indices and the box."""

    def __init__(self, pos, mat=None, materials=None, grid_unit=1.0, shape=None):
        self.shape = shape  # the generating shape ("sphere", ...), None for a dipole list
        self.pos = np.ascontiguousarray(pos, np.uint32).reshape(-1, 3)
        self.nv = self.pos.shape[0]
        self.mat = np.zeros(self.nv, np.uint8) if mat is None else np.ascontiguousarray(mat, np.uint8)
        self.materials = [complex(1)] if materials is None else list(materials)
        self.grid_unit = float(grid_unit)
        self.box = tuple(int(v) + 1 for v in self.pos.max(axis=0)) if self.nv else (0, 0, 0)

    @property
    def origin(self):
        """Position of grid point (0,0,0): DipoleGeometry::moveToCenter."""
        return -np.array([(b - 1.0) / 2 for b in self.box]) * self.grid_unit

    def coords(self):
        return self.origin[None, :] + self.pos.astype(np.float64) * self.grid_unit


def _zyx_grid(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def sphere(radius_units, material=1.5 + 0.1j, grid_unit=1.0):
    """Shapes.cpp:53-73; radius in grid units (radius / gridUnit)."""
    r = float(radius_units)
    n = int(math.ceil(2 * r))
    x, y, z = _zyx_grid(n, n, n)
    xd, yd, zd = x - r + 0.5, y - r + 0.5, z - r + 0.5
    keep = xd * xd + yd * yd + zd * zd <= r * r
    pos = np.stack([x[keep], y[keep], z[keep]], axis=1)
    return DipoleGeometry(pos, None, [material], grid_unit, shape="sphere")


def box(nx, ny, nz, material=1.5 + 0.1j, grid_unit=1.0):
    """Shapes.cpp:130-143."""
    x, y, z = _zyx_grid(int(math.ceil(nx)), int(math.ceil(ny)), int(math.ceil(nz)))
    return DipoleGeometry(np.stack([x, y, z], axis=1), None, [material], grid_unit, shape="box")


def cylinder(length_units, radius_units, material=1.5 + 0.1j, grid_unit=1.0):
    """Shapes.cpp:88-106 (axis along x)."""
    r = float(radius_units)
    n = int(math.ceil(2 * r))
    L = int(math.ceil(length_units))
    x, y, z = _zyx_grid(L, n, n)
    yd, zd = y - r + 0.5, z - r + 0.5
    keep = yd * yd + zd * zd <= r * r
    return DipoleGeometry(np.stack([x[keep], y[keep], z[keep]], axis=1), None, [material], grid_unit,
                          shape="cylinder")


def ellipsoid(a, b, c, material=1.5 + 0.1j, grid_unit=1.0):
    """Semi-axes (a, b, c) in grid units, sphere centre rule per axis."""
    nx, ny, nz = (int(math.ceil(2 * v)) for v in (a, b, c))
    x, y, z = _zyx_grid(nx, ny, nz)
    xd, yd, zd = (x - a + 0.5) / a, (y - b + 0.5) / b, (z - c + 0.5) / c
    keep = xd * xd + yd * yd + zd * zd <= 1.0
    return DipoleGeometry(np.stack([x[keep], y[keep], z[keep]], axis=1), None, [material], grid_unit,
                          shape="ellipsoid")


def load_dipole_file(text, materials=None, grid_unit=1.0):
    """ADDA dipole list (Load.cpp:37-69): '#' comments, 'x y z [mat]', mat 1-based."""
    rows = []
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("#") or line.startswith("N"):
            continue
        v = [int(t) for t in line.split()]
        rows.append((v[0], v[1], v[2], (v[3] if len(v) > 3 else 1) - 1))
    a = np.array(rows, np.int64)
    nmat = int(a[:, 3].max()) + 1
    mats = materials if materials is not None else [1.5 + 0.1j] * nmat
    return DipoleGeometry(a[:, :3], a[:, 3], mats, grid_unit)


class DDAParams:
    """Grid, stride and wave number of one problem (DDA/DDAParams.cpp:93-214)."""

    def __init__(self, geometry, lambda_, grid=(0, 0, 0), support_non_pot=True):
        self.geometry = geometry
        self.lambda_ = float(lambda_)
        self.grid = tuple(fft_fit(2 * b, g, support_non_pot) for b, g in zip(geometry.box, grid))
        self.nv = geometry.nv
        self.vec_stride = (self.nv + 15) // 16 * 16
        self.wave_num = 2 * math.pi / self.lambda_
        self.kd = self.wave_num * geometry.grid_unit

    @property
    def vec_size(self):
        return 3 * self.vec_stride

    def to_soa(self, v3):
        """[nv, 3] -> SoA [3 * vec_stride] with zero padding."""
        out = np.zeros(3 * self.vec_stride, np.complex128)
        for c in range(3):
            out[c * self.vec_stride: c * self.vec_stride + self.nv] = v3[:, c]
        return out

    def from_soa(self, v):
        return np.stack([v[c * self.vec_stride: c * self.vec_stride + self.nv] for c in range(3)], axis=1)


def beam_pols(prop):
    """Beam.cpp:34-51: returns (prop, pol1 'y polarisation', pol2 'x polarisation')."""
    p = np.asarray(prop, np.float64)
    p = p / np.linalg.norm(p)
    if abs(p[2]) >= 1 - 1e-6:
        px = np.array([p[2], 0.0, 0.0])
        py = np.array([0.0, 1.0, 0.0])
    else:
        t = math.sqrt(1 - p[2] * p[2])
        px = np.array([p[0] * p[2] / t, p[1] * p[2] / t, -t])
        py = np.array([-p[1] / t, p[0] / t, 0.0])
    return p, py, px


def couple_constants(params, prop, incpol, kind="ldr"):
    """(cc, cc_sqrt), each [nmat, 3] complex in m^3 and m^(3/2).

    CM: 3 V/(4 pi) (m^2 - 1)/(m^2 + 2); LDR divides by 1 - (a/V) t1 with
    t1 = (b1 + (b2 + b3 S) m^2) kd^2 + i 2/3 kd^3, S = sum prop_i^2 pol_i^2
    ("ldr-avgpol": S = (1 - sum prop_i^4)/2)."""
    g = params.geometry
    vol = g.grid_unit ** 3
    kd = params.kd
    mats = np.array([np.broadcast_to(np.asarray(m, np.complex128), (3,)) for m in g.materials])
    m2 = mats * mats
    unc = 3 * vol / (4 * math.pi) * (m2 - 1) / (m2 + 2)
    if kind == "cm":
        cc = unc
    else:
        p2 = np.asarray(prop, np.float64) ** 2
        if kind == "ldr-avgpol":
            S = 0.5 * (1 - np.sum(p2 * p2))
        else:
            S = float(np.sum(p2 * np.asarray(incpol, np.float64) ** 2))
        t1 = (LDR_B1 + (LDR_B2 + LDR_B3 * S) * m2) * kd * kd + 1j * (2.0 / 3.0) * kd ** 3
        cc = unc / (1 - (unc / vol) * t1)
    return cc, np.sqrt(cc)


def rotation_symmetric(geometry, prop, orientation_none=True):
    """Whether one solve (polarisation 1) covers both polarisations: the
    reference's rule (DDA/DDAMain.cpp:105-121, 252-260) for a plane wave
    (Beam.cpp:107-109): propagation along +z (|prop - z|^2 <= 1e-12), no
    particle rotation, every material isotropic and a shape whose dipole model
    is symmetric under a quarter turn about z (the sphere alone,
    Shapes.hpp:48). Polarisation 2's solution is then polarisation 1's rotated
    (FarFieldCalc.cpp:167-168) and its cross sections are polarisation 1's
    (DDAMain.cpp:461-464)."""
    p = np.asarray(prop, np.float64)
    p = p / np.linalg.norm(p)
    if float(np.sum((p - np.array([0.0, 0.0, 1.0])) ** 2)) > 1e-12 or not orientation_none:
        return False
    for m in geometry.materials:
        v = np.broadcast_to(np.asarray(m, np.complex128), (3,))
        if not (v[0] == v[1] and v[0] == v[2]):
            return False
    return geometry.shape == "sphere"


def einc_plane(params, prop, pol):
    """Plane wave E_i = exp(i k r_i . prop) pol (Beam.cpp:110-118), SoA."""
    r = params.geometry.coords()
    ph = np.exp(1j * params.wave_num * (r @ np.asarray(prop, np.float64)))
    return params.to_soa(ph[:, None] * np.asarray(pol, np.float64)[None, :])


def directions(n, seed):
    """n propagation directions uniform on the unit sphere (z uniform in [-1, 1],
    azimuth uniform in [0, 2 pi)), reproducible from the seed."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2 * math.pi, n)
    t = np.sqrt(1 - z * z)
    return np.stack([t * np.cos(phi), t * np.sin(phi), z], axis=1)


def incident_batch(params, dirs, kind="ldr"):
    """Right-hand sides of an orientation batch: for each direction both
    polarisations of Beam.cpp:34-51, in that order. Returns (einc [2n, vec_size],
    cc_sqrt [2n, nmat, 3], props [2n, 3], pols [2n, 3])."""
    eincs, ccs, props, pols = [], [], [], []
    for d in dirs:
        prop, pol1, pol2 = beam_pols(d)
        for pol in (pol1, pol2):
            eincs.append(einc_plane(params, prop, pol))
            ccs.append(couple_constants(params, prop, pol, kind)[1])
            props.append(prop)
            pols.append(pol)
    return np.array(eincs), np.array(ccs), np.array(props), np.array(pols)


def ext_cross(params, pvec, einc):
    """Cext = 4 pi k sum Im(p_i . conj(E_i)) (Beam.cpp:58-68)."""
    p = params.from_soa(pvec)
    e = params.from_soa(einc)
    return 4 * math.pi * params.wave_num * float(np.sum(np.imag(np.sum(p * np.conj(e), axis=1))))


def abs_cross(params, pvec, cc):
    """Cabs = 4 pi k sum (-Im(1/a) - 2k^3/3) |p|^2 (AbsCross.cpp:27-48)."""
    k = params.wave_num
    mult = -np.imag(1.0 / np.asarray(cc, np.complex128).reshape(-1, 3)) - 2.0 * k ** 3 / 3.0
    p = params.from_soa(pvec)
    m = params.geometry.mat
    return 4 * math.pi * k * float(np.sum(mult[m] * np.abs(p) ** 2))


def far_field(params, pvec, n):
    """Scattered amplitude in direction n (CpuFieldCalculator.cpp:48-98)."""
    g = params.geometry
    n = np.asarray(n, np.float64)
    p = params.from_soa(pvec)
    ph = np.exp(-1j * params.kd * (g.pos.astype(np.float64) @ n))
    s = (p * ph[:, None]).sum(axis=0)
    t = s - n * (n @ s)
    return 1j * params.wave_num ** 2 * np.exp(-1j * params.wave_num * (g.origin @ n)) * t


def ext_cross_forward(params, pvec, prop, pol):
    """Cext from the forward amplitude (Beam.cpp:125-133)."""
    e = far_field(params, pvec, prop)
    return float(-np.real(e) @ np.asarray(pol, np.float64)) * 4 * math.pi / params.wave_num


def random_dipole_vector(params, seed=0x5EED, dtype=np.complex128):
    """x with nv entries per component, re, im ~ U[-1, 1), padding zero."""
    rng = np.random.default_rng(seed)
    v = rng.uniform(-1, 1, (params.nv, 3)) + 1j * rng.uniform(-1, 1, (params.nv, 3))
    return params.to_soa(v).astype(dtype)
