"""Host-side pieces of the library that run without a GPU."""
import math
import os

import numpy as np

import dda_amd as D
from dda_amd import problem as P

import cases


def test_couple_constants_match_host():
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    for kind in ("cm", "ldr", "ldr-avgpol"):
        c_ref, s_ref = P.couple_constants(params, prop, pol, kind)
        c, s = D.couple_constants(g.materials, g.grid_unit, params.kd, prop, pol, kind)
        np.testing.assert_allclose(c, c_ref, rtol=1e-13)
        np.testing.assert_allclose(s, s_ref, rtol=1e-13)
        # and the oracle's restatement
        import oracle as O
        oc, os_ = O.couple_constants(kind, g.materials, g.grid_unit, params.kd, prop, pol)
        np.testing.assert_allclose(c, oc, rtol=1e-12)


def test_dippol_and_jones_text_round_trip(tmp_path):
    from dda_amd import output as O
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    rng = np.random.default_rng(1)
    p = P.random_dipole_vector(params, seed=2) * 1e-20
    f = tmp_path / "DipPol-Pol1.txt"
    O.write_dippol_txt(f, params, p, 1)
    head = open(f).readline().split()
    assert head[:5] == ["#x", "y", "z", "|DipolePolarization-Pol1|^2", "DipolePolarization-Pol1.x.r"]
    back = O.read_dippol_txt(f, params)
    assert O.field_diff(back, p) < 1e-9  # 10 significant digits
    J = rng.normal(size=(12, 2, 2)) + 1j * rng.normal(size=(12, 2, 2))
    th = np.radians(np.repeat(np.arange(0, 180, 60.0), 4))
    ph = np.radians(np.tile(np.arange(0, 360, 90.0), 3))
    fj = tmp_path / "jones.txt"
    O.write_jones_txt(fj, th, ph, J)
    t2, p2, J2 = O.read_jones_txt(fj)
    assert np.allclose(t2, np.degrees(th)) and np.allclose(p2, np.degrees(ph))
    assert O.field_diff(J2, J) < 1e-9
    assert O.field_diff(J, J) == 0.0 and abs(O.field_diff(J, 2 * J) - math.sqrt(0.4)) < 1e-12


def test_fft_fit_static():
    """The fast-grid rule: smallest compile-time planned length >= x, else the smooth fftFit."""
    want = {20: 32, 32: 32, 33: 48, 50: 64, 97: 128, 100: 128, 130: 192, 200: 256, 300: 512, 600: 1024,
            1100: 1120}
    for x, n in want.items():
        assert D.lib().dda_fft_fit_static(x) == n, x


def test_rotation_symmetric_rule():
    """DDAMain.cpp:105-121, 252-260: one solve covers both polarisations only
    for an isotropic sphere lit along +z without rotation."""
    s = P.sphere(5.0, 1.5 + 0.1j, 1e-7)
    assert P.rotation_symmetric(s, (0, 0, 1))
    assert P.rotation_symmetric(s, (0, 0, 2.0))  # normalised first
    assert not P.rotation_symmetric(s, (0, 0, -1))
    assert not P.rotation_symmetric(s, (0.01, 0, 1))
    assert not P.rotation_symmetric(s, (0, 0, 1), orientation_none=False)
    assert not P.rotation_symmetric(P.box(6, 6, 6, 1.5, 1e-7), (0, 0, 1))
    assert not P.rotation_symmetric(P.ellipsoid(4, 4, 4, 1.5, 1e-7), (0, 0, 1))
    aniso = P.sphere(5.0, 1.5 + 0.1j, 1e-7)
    aniso.materials = [np.array([1.5, 1.5, 1.6 + 0.1j])]
    assert not P.rotation_symmetric(aniso, (0, 0, 1))
    iso3 = P.sphere(5.0, 1.5 + 0.1j, 1e-7)
    iso3.materials = [np.array([1.5 + 0.1j] * 3)]
    assert P.rotation_symmetric(iso3, (0, 0, 1))
    with open(cases.os.path.join(cases.GOLDEN, "box_10_8_6.geom")) as f:
        assert not P.rotation_symmetric(P.load_dipole_file(f.read(), [1.33], 1e-7), (0, 0, 1))


def test_anisotropic_couple_constants_per_axis():
    """dda_couple_constants_diag (DiagMatrix3 materials,
    PolarizabilityDescription.cpp:69-75, 90-96): axis c of an anisotropic
    material equals the isotropic result for that axis's index, and the host
    twin agrees."""
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    mats = [np.array([1.5 + 0.1j, 1.7 + 0.02j, 2.0 + 0.3j]), 1.33 + 0.0j]
    for kind in ("cm", "ldr", "ldr-avgpol"):
        c, s = D.couple_constants(mats, g.grid_unit, params.kd, prop, pol, kind)
        assert c.shape == (2, 3)
        for ax in range(3):
            ci, si = D.couple_constants([mats[0][ax], mats[1]], g.grid_unit, params.kd, prop, pol, kind)
            np.testing.assert_array_equal(c[:, ax], ci[:, ax])
            np.testing.assert_array_equal(s[:, ax], si[:, ax])
        g2 = P.DipoleGeometry(g.pos, g.mat, mats, g.grid_unit)
        c_ref, s_ref = P.couple_constants(P.DDAParams(g2, params.lambda_, params.grid), prop, pol, kind)
        np.testing.assert_allclose(c, c_ref, rtol=1e-13)
        np.testing.assert_allclose(s, s_ref, rtol=1e-13)


def test_hdf5_dippol_and_jones_files(tmp_path):
    """DipPol and Jones files in the reference's MATLAB 7.3 layout
    (HDF5::matlabSerialize, EMSim/DataFilesDDA.cpp:58-63,
    EMSim/JonesCalculus.hpp:82-93): values read back bit for bit, the "MAT"
    user block, and -- through the HDF5 tools' own reader, h5dump -- the
    struct groups, classes, string encoding, dims and the Fortran-order Jones
    array."""
    import shutil
    import subprocess
    from dda_amd import hdf5 as H
    if not H.available():
        import pytest
        pytest.skip("no HDF5 C library")
    g, params, cc, cs, prop, pol = cases.build("two_materials")
    _, p1, p2 = P.beam_pols(prop)
    pv = P.random_dipole_vector(params, seed=4)
    f1 = tmp_path / "DipPol-Pol1.hdf5"
    H.write_dippol(f1, params, pv, 1, prop, [p1, p2], geometry_string="sphere", cmdline="dda")
    assert np.array_equal(H.read_dippol(f1, params), pv)
    head = open(f1, "rb").read(512)
    assert head[:4] == b"MAT\x00" and head[125:128] == b"\x02IM"
    assert np.array_equal(H.read_dataset(f1, "/Geometry/DipolePositions", dtype=np.uint32), g.pos)
    assert np.array_equal(H.read_dataset(f1, "/Geometry/Size", dtype=np.uint32), np.array(g.box))
    np.testing.assert_allclose(H.read_dataset(f1, "/Parameters/MethodParameters/Frequency"),
                               [299792458.0 / params.lambda_], rtol=0)
    rng = np.random.default_rng(5)
    n = 7
    J = rng.normal(size=(n, 2, 2)) + 1j * rng.normal(size=(n, 2, 2))
    th, ph = np.linspace(0, 1, n), np.linspace(0, 2, n)
    f2 = tmp_path / "Far.hdf5"
    H.write_jones(f2, params, th, ph, J, prop, [p1, p2])
    d = H.read_dataset(f2, "/JonesFarField/Data", complex_=True)
    assert d.shape == (1, n, 2, 2)
    for k in range(n):
        for i in range(2):
            for j in range(2):
                assert d[0, k, j, i] == J[k, i, j]  # Data[i][j][k][0], Fortran order
    np.testing.assert_array_equal(H.read_dataset(f2, "/JonesFarField/Theta"), th)
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if not os.path.exists(h5dump):
        return
    txt = subprocess.run([h5dump, "-A", str(f1)], capture_output=True, text=True, check=True).stdout
    assert 'GROUP "Field"' in txt and 'GROUP "Geometry"' in txt and 'GROUP "Parameters"' in txt
    assert '(0): "struct"' in txt and '(0): "char"' in txt and '(0): "uint32"' in txt
    assert 'DATASET "Periodicity"' in txt and "SIMPLE { ( 0, 3 ) / ( 0, 3 ) }" in txt
    assert 'H5T_IEEE_F64LE "real";' in txt and 'H5T_IEEE_F64LE "imag";' in txt
    assert "SIMPLE { ( %d, 3 ) / ( %d, 3 ) }" % (params.nv, params.nv) in txt
    t2 = subprocess.run([h5dump, "-A", str(f2)], capture_output=True, text=True, check=True).stdout
    assert 'DATASET "DipolePositions"' not in t2  # far-field files carry the geometry without positions
    assert "SIMPLE { ( 1, %d, 2, 2 ) / ( 1, %d, 2, 2 ) }" % (n, n) in t2
    # the Type string as UTF-16 code units, dims (len, 1)
    t3 = subprocess.run([h5dump, "-d", "/Type", str(f2)], capture_output=True, text=True, check=True).stdout
    assert "SIMPLE { ( 13, 1 ) / ( 13, 1 ) }" in t3
    assert all("(%d,0): %d" % (i, ord(c)) in t3 for i, c in enumerate("JonesFarField"))
