"""Text output formats of the reference for results of this path.

DipPol (dipole polarisation) files follow DDAFieldFile::writeTxt
(EMSim/DataFilesDDA.cpp:34-56) and are read back as Load::loadFields does
(DDA/Load.cpp:76-107); Jones files follow JonesCalculus::storeTxt
(EMSim/JonesCalculus.cpp:88-116); field_diff is FieldDiff's normalised RMS
(DDA/FieldDiff.cpp:190-192). HDF5 output (the reference's default) needs h5py,
which this environment lacks, so only the text forms are written.
"""
import math

import numpy as np


def _g10(v):
    # C++ iostream with precision(10) and default float field: %.10g
    return "%.10g" % v


def write_dippol_txt(path, params, pvec, polarization=1, name="DipolePolarization"):
    """One line per dipole: x y z |P|^2 P.x.r P.x.i P.y.r P.y.i P.z.r P.z.i."""
    n = name + ("-Pol%d" % polarization if polarization else "")
    coords = params.geometry.coords()
    p = params.from_soa(np.asarray(pvec))
    with open(path, "w") as f:
        f.write("#x y z |%s|^2 %s.x.r %s.x.i %s.y.r %s.y.i %s.z.r %s.z.i\n" % ((n,) * 7))
        for c, v in zip(coords, p):
            vals = [c[0], c[1], c[2], float(np.sum(np.abs(v) ** 2)),
                    v[0].real, v[0].imag, v[1].real, v[1].imag, v[2].real, v[2].imag]
            f.write(" ".join(_g10(x) for x in vals) + "\n")


def read_dippol_txt(path, params):
    """The SoA vector [3 * vec_stride] stored by write_dippol_txt, checking the
    coordinates and |P|^2 against the geometry as Load::loadFields does."""
    rows = np.loadtxt(path, comments="#", ndmin=2)
    if rows.shape != (params.nv, 10):
        raise ValueError("%s: expected %d rows of 10 columns, got %s" % (path, params.nv, rows.shape))
    coords = params.geometry.coords()
    if np.max(np.sum((rows[:, :3] - coords) ** 2, axis=1)) >= 1e-5:
        raise ValueError("%s: dipole coordinates do not match the geometry" % path)
    v = rows[:, 4:10:2] + 1j * rows[:, 5:10:2]
    if np.max(np.abs(np.sum(np.abs(v) ** 2, axis=1) - rows[:, 3])) >= 1e-5:
        raise ValueError("%s: |P|^2 column does not match the components" % path)
    return params.to_soa(v)


def write_jones_txt(path, theta, phi, J, store_phi=True):
    """theta, phi in radians [count]; J [count, 2, 2] complex. Columns:
    theta [phi] s1 s2 s3 s4 (real, imag) with s1 = J[1][1], s2 = J[0][0],
    s3 = J[0][1], s4 = J[1][0]."""
    J = np.asarray(J)
    with open(path, "w") as f:
        f.write("#theta " + ("phi " if store_phi else "") +
                "s1.r s1.i s2.r s2.i s3.r s3.i s4.r s4.i\n")
        for t, p, m in zip(theta, phi, J):
            s = [m[1, 1], m[0, 0], m[0, 1], m[1, 0]]
            head = "%.2f " % math.degrees(t) + ("%.2f " % math.degrees(p) if store_phi else "")
            f.write(head + " ".join("%.10e %.10e" % (x.real, x.imag) for x in s) + "\n")


def read_jones_txt(path, store_phi=True):
    """-> (theta_deg, phi_deg or None, J [count, 2, 2])."""
    rows = np.loadtxt(path, comments="#", ndmin=2)
    off = 2 if store_phi else 1
    s = rows[:, off::2] + 1j * rows[:, off + 1::2]
    J = np.zeros((rows.shape[0], 2, 2), np.complex128)
    J[:, 1, 1], J[:, 0, 0], J[:, 0, 1], J[:, 1, 0] = s[:, 0], s[:, 1], s[:, 2], s[:, 3]
    return rows[:, 0], (rows[:, 1] if store_phi else None), J


def field_diff(a, b):
    """sqrt(sum |a - b|^2 / ((sum |a|^2 + sum |b|^2) / 2)) over all values."""
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    d = np.sum(np.abs(a - b) ** 2)
    return float(math.sqrt(d / ((np.sum(np.abs(a) ** 2) + np.sum(np.abs(b) ** 2)) / 2)))


def angle_grid(theta_step_deg=2.0, phi_step_deg=6.0):
    """theta 0..180 and phi 0..(360 - step) on a regular grid (the Jones grid of the 40-across sphere case
    is 2 deg x 6 deg) -> (theta0, dtheta, ntheta, phi0, dphi, nphi) in radians."""
    nt = int(round(180.0 / theta_step_deg)) + 1
    npf = int(round(360.0 / phi_step_deg))
    return 0.0, math.radians(theta_step_deg), nt, 0.0, math.radians(phi_step_deg), npf
