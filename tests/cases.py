"""Seeded parity cases shared by the GPU tests: geometry, wavelength, materials.

Every case is built twice from the same numbers: once through the product's
set-up (dda_amd.problem) for the HIP path, once through the oracle.
"""
import math
import os

import numpy as np

from dda_amd import problem as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pinwheel_positions():
    """A chiral dipole set: nine dipoles of a twisted arm, repeated by quarter
    turns about the z axis through the centre of an 8x8x3 box, sorted by
    (z, y, x). It is symmetric under those quarter turns and under no mirror
    plane."""
    B = 8
    arm = [(4, 3, 0), (5, 3, 0), (6, 3, 0), (7, 3, 0), (7, 4, 1), (7, 5, 1), (6, 6, 2), (4, 4, 0), (5, 4, 1)]
    pts = set()
    for (x, y, z) in arm:
        for _ in range(4):
            pts.add((x, y, z))
            x, y = B - 1 - y, x
    return np.array(sorted(pts, key=lambda p: (p[2], p[1], p[0])), np.uint32)


def _geom(name):
    d = 0.5e-6
    if name == "pinwheel":  # chiral, quarter-turn symmetric about z (32 dipoles)
        return P.DipoleGeometry(pinwheel_positions(), None, [1.5 + 0.1j], d), 3e-6
    if name == "sphere_r4":
        return P.sphere(4.0, 1.5 + 0.1j, d), 6e-6
    if name == "sphere_m3_r8":  # high index: erratic residuals (stagnation tests)
        return P.sphere(8.0, 3.0 + 0.01j, d), 6e-6
    if name == "sphere_r7":
        return P.sphere(7.0, 1.5 + 0.1j, d), 6e-6
    if name == "box_10_8_6":
        with open(os.path.join(GOLDEN, "box_10_8_6.geom")) as f:
            g = P.load_dipole_file(f.read(), [1.33 + 0.01j], d)
        return g, 2 * math.pi * d * 1.33 / 0.6  # |m| kd = 0.6
    if name == "ellipsoid_5_4_3":
        return P.ellipsoid(5, 4, 3, 1.33 + 0.01j, d), 6e-6
    if name == "two_materials":
        g = P.sphere(5.0, 1.5 + 0.1j, d)
        c = g.pos.astype(np.float64) - (np.array(g.box) - 1) / 2.0
        g.mat = (np.sum(c * c, axis=1) < 2.5 ** 2).astype(np.uint8)
        g.materials = [1.5 + 0.1j, 2.0 + 0.3j]
        return g, 6e-6
    if name == "cylinder_9_3":
        return P.cylinder(9, 3, 1.7 + 0.05j, d), 5e-6
    if name == "sphere_c1":  # 40 dipoles across, lambda 6 um, d 0.5 um (33,552 dipoles)
        return P.sphere(20.0, 1.5 + 0.1j, d), 6e-6
    if name == "sphere_c2":  # 128 dipoles across, |m| kd = 0.5 (1,099,136 dipoles, 256^3 grid)
        m = 1.5 + 0.1j
        return P.sphere(64.0, m, d), 2 * math.pi * d * abs(m) / 0.5
    raise KeyError(name)


SMALL_CASES = ["sphere_r4", "box_10_8_6", "ellipsoid_5_4_3", "two_materials", "cylinder_9_3"]


def build(name, grid=(0, 0, 0), support_non_pot=True, prop=(0, 0, 1), pol_index=1, kind="ldr"):
    g, lam = _geom(name)
    params = P.DDAParams(g, lam, grid, support_non_pot)
    prop_n, pol1, pol2 = P.beam_pols(prop)
    pol = pol1 if pol_index == 1 else pol2
    cc, cs = P.couple_constants(params, prop_n, pol, kind)
    return g, params, cc, cs, prop_n, pol


def oracle_problem(g, params, cs, dmat_precision=80):
    import oracle as O
    return O.Problem(g.pos, g.mat, g.box, params.grid, cs, g.grid_unit, params.wave_num,
                     dmat_precision=dmat_precision)


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
