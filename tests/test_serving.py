"""The seeded serving job stream (CPU): the distributions it draws from, the
incident frame of each job, and the arrival process."""
import math

import numpy as np

from dda_amd import problem as P
from dda_amd import serving as S


def test_stream_is_seeded_and_within_its_ranges():
    a = S.make_stream(8, seed=42)
    b = S.make_stream(8, seed=42)
    for ja, jb in zip(a, b):
        assert ja.shape == jb.shape and ja.dims == jb.dims and ja.m == jb.m
        assert np.array_equal(ja.einc, jb.einc)
    for j in a:
        assert j.shape in ("sphere", "box", "ellipsoid")
        assert 16 <= j.dims[0] <= 96
        assert all(max(1, round(0.5 * j.dims[0])) <= o <= j.dims[0] for o in j.dims[1:])
        assert 1.2 <= j.m.real <= 2.0 and 0.0 <= j.m.imag <= 0.2
        # |m| kd = 0.5 with kd = wave_num * grid unit
        assert abs(abs(j.m) * j.params.wave_num * S.GRID_UNIT - 0.5) < 1e-12
        assert j.params.nv == len(j.geometry.pos) > 0
        assert j.einc.shape == (2, 3 * j.params.vec_stride)
        assert j.cc.shape == j.cc_sqrt.shape == (2, len(j.geometry.materials), 3)
        np.testing.assert_allclose(j.cc_sqrt ** 2, j.cc, rtol=1e-12)


def test_job_frame_is_a_rotation_of_the_beam():
    """prop and the two polarisations are the rotated z-beam frame: orthonormal,
    right-handed like the unrotated one."""
    rng = np.random.default_rng(3)
    _, q1, q2 = P.beam_pols((0.0, 0.0, 1.0))
    hand = np.dot(np.cross(q1, q2), (0.0, 0.0, 1.0))
    for _ in range(5):
        j = S.make_job(rng, dim=16)
        k, p1, p2 = j.prop, j.pols[0], j.pols[1]
        for v in (k, p1, p2):
            assert abs(np.linalg.norm(v) - 1) < 1e-12
        assert max(abs(np.dot(k, p1)), abs(np.dot(k, p2)), abs(np.dot(p1, p2))) < 1e-12
        assert abs(np.dot(np.cross(p1, p2), k) - hand) < 1e-12


def test_incident_field_is_the_rotated_plane_wave():
    rng = np.random.default_rng(5)
    j = S.make_job(rng, shape="box", dim=16, other=(8, 8), m=1.5 + 0.05j)
    g, params = j.geometry, j.params
    origin = -(np.array(g.box, float) - 1) / 2 * g.grid_unit
    r = origin + g.pos.astype(float) * g.grid_unit
    phase = np.exp(1j * params.wave_num * (r @ j.prop))
    S_ = params.vec_stride
    for p in range(2):
        for c in range(3):
            np.testing.assert_allclose(j.einc[p, c * S_:c * S_ + params.nv], j.pols[p][c] * phase, atol=1e-12)
            assert np.all(j.einc[p, c * S_ + params.nv:(c + 1) * S_] == 0)


def test_poisson_arrivals():
    jobs = S.make_stream(6, seed=7, rate=50.0)
    t = [j.arrival for j in jobs]
    assert t[0] > 0 and all(b > a for a, b in zip(t, t[1:]))
    burst = S.make_stream(3, seed=7)
    assert all(j.arrival == 0.0 for j in burst)
    assert math.isfinite(sum(t))
