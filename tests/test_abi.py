"""CPU-side checks of the product: the C ABI loads and exports every declared
symbol, and the host-side problem set-up agrees with the oracle's restatement.
No GPU call is made here."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import dda_amd as D
from dda_amd import problem as P
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dda_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dda_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = D.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(D.EXPORTS) == names


def test_fft_fit_matches_oracle():
    for x in list(range(2, 600, 2)) + [1022, 1024, 2048]:
        assert D.fft_fit(x) == O.fft_fit(x)
        assert D.fft_fit(x, 0, False) == O.fft_fit(x, 0, False)
    assert D.fft_fit(80, 96) == 96 and D.fft_fit(100, 96) == 0


@pytest.mark.parametrize("shape,args,oargs", [
    ("sphere", (20.0,), (20.0, 0, 0)),
    ("sphere", (64.0,), (64.0, 0, 0)),
    ("box", (10, 8, 6), (10, 8, 6)),
    ("ellipsoid", (32, 24, 16), (32, 24, 16)),
    ("cylinder", (9, 3), (9, 3, 0)),
])
def test_geometry_matches_oracle(shape, args, oargs):
    g = getattr(P, shape)(*args)
    pos, box = O.geometry(shape, *oargs)
    np.testing.assert_array_equal(g.pos, pos)
    assert tuple(g.box) == tuple(int(b) for b in box)


def test_sphere_c2_count():
    assert P.sphere(64.0).nv == 1099136


@pytest.mark.parametrize("prop", [(0, 0, 1), (0.3, -0.5, 0.8), (1, 0, 0)])
@pytest.mark.parametrize("kind", ["cm", "ldr", "ldr-avgpol"])
def test_couple_constants_and_beam_match_oracle(prop, kind):
    g = P.sphere(3.0, 1.5 + 0.1j, 0.5e-6)
    params = P.DDAParams(g, 6e-6)
    p, pol1, pol2 = P.beam_pols(prop)
    op, opol1, opol2 = O.beam_pols(prop)
    np.testing.assert_allclose(p, op, atol=1e-15)
    np.testing.assert_allclose(pol1, opol1, atol=1e-15)
    np.testing.assert_allclose(pol2, opol2, atol=1e-15)
    for pol in (pol1, pol2):
        cc, cs = P.couple_constants(params, p, pol, kind)
        occ, ocs = O.couple_constants(kind, 1.5 + 0.1j, g.grid_unit, params.kd, op, pol)
        np.testing.assert_allclose(cc, occ, rtol=1e-13)
        np.testing.assert_allclose(cs, ocs, rtol=1e-13)
        e = P.einc_plane(params, p, pol)
        oe = O.einc_plane(g.pos, g.box, params.vec_stride, g.grid_unit, params.wave_num, op, pol)
        np.testing.assert_allclose(e, oe, atol=1e-13)


def test_observables_match_oracle():
    g = P.sphere(3.0, 1.5 + 0.1j, 0.5e-6)
    params = P.DDAParams(g, 6e-6)
    prop, pol1, _ = P.beam_pols((0, 0, 1))
    cc, cs = P.couple_constants(params, prop, pol1)
    orc = O.Problem(g.pos, g.mat, g.box, params.grid, cs, g.grid_unit, params.wave_num)
    e = P.einc_plane(params, prop, pol1)
    st, p, it, _ = orc.solve("qmr", e, 1e-10)
    assert abs(P.ext_cross(params, p, e) / orc.ext_cross_sum(p, e) - 1) < 1e-12
    assert abs(P.abs_cross(params, p, cc[0]) / orc.abs_cross(p, cc) - 1) < 1e-12
    assert abs(P.ext_cross_forward(params, p, prop, pol1) / orc.ext_cross_forward(p, prop, pol1) - 1) < 1e-12
    n = np.array([0.3, 0.4, math.sqrt(1 - 0.25)])
    np.testing.assert_allclose(P.far_field(params, p, n), orc.calc_field(p, n), rtol=1e-12)


def test_load_dipole_file_fixture():
    path = os.path.join(ROOT, "tests", "golden", "box_10_8_6.geom")
    g = P.load_dipole_file(open(path).read())
    assert g.nv == 480 and g.box == (10, 8, 6)


def test_static_z_pass_only_where_its_offsets_fit_32_bits():
    """The static Z pass addresses the tensor octant and Gz/8 planes of the
    grid with 32-bit byte offsets; grids beyond that fall back to the run-time
    planned pass (64-bit addressing) instead of wrapping."""
    assert D.grid_plan((256, 256, 256), 32) == (True, True)
    assert D.grid_plan((256, 256, 256), 64) == (True, True)
    # 1024^3 in FP64: octant 513^3 * 16 B = 2.16 GB, 129 planes * 16 MiB = 2.16 GB
    assert D.grid_plan((1024, 1024, 1024), 64) == (True, True)
    # box 672 x 784 x 512: octant 673 * 785 * 513 * 16 B = 4.34 GB > 4 GiB in FP64
    assert D.grid_plan((1344, 1568, 1024), 64) == (False, False)
    assert D.grid_plan((1344, 1568, 1024), 32) == (False, True)
    # no static plan for 80 (C1's grid): run-time planned either way
    assert D.grid_plan((80, 80, 80), 64) == (False, False)
