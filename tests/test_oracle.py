"""Pins the CPU oracle against what mathematics and library routines fix.

Each check uses a different computation from the oracle's own: brute-force DFT
and numpy.fft for the FFT, finite differences of the scalar Green function for
the dyadic tensor, a dense direct sum for the FFT matvec, numpy.linalg.solve for
the solvers, the optical theorem for Cext, and limits of Mie and LDR theory.
"""
import math

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("n", [8, 12, 14, 18, 30, 60, 80, 96, 128, 192, 256])
def test_fft_against_brute_force_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    k = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(k, k) / n)  # direct definition
    ref = x @ W.T
    f = O.fft(x)
    assert np.max(np.abs(f - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert np.max(np.abs(f - np.fft.fft(x))) <= 1e-13 * np.max(np.abs(ref))
    # inverse is exp(+...), unnormalised: round trip gives n x
    assert np.max(np.abs(O.fft(f, inverse=True) - n * x)) <= 1e-12 * n * np.max(np.abs(x))
    # Parseval
    assert abs(np.sum(np.abs(f) ** 2) - n * np.sum(np.abs(x) ** 2)) <= 1e-12 * n * np.sum(np.abs(x) ** 2)


def test_fft_delta_and_constant():
    n = 80
    d = np.zeros(n, complex)
    d[0] = 1
    np.testing.assert_allclose(O.fft(d), np.ones(n), atol=1e-15)
    np.testing.assert_allclose(O.fft(np.ones(n, complex)), d * n, atol=1e-12)


def test_fft_fit_rules():
    # smooth rule (DDAParams.cpp:63-89): even 2^a 3^b 5^c 7^d >= x
    assert [O.fft_fit(x) for x in (80, 81, 256, 22, 26)] == [80, 84, 256, 24, 28]
    # power-of-two rule (:54-61) and override (:46-49)
    assert [O.fft_fit(x, 0, False) for x in (80, 96, 256, 257)] == [128, 128, 256, 512]
    assert O.fft_fit(80, 96) == 96


def _scalar_green(r, k):
    rr = np.linalg.norm(r)
    return np.exp(1j * k * rr) / rr


def test_green_tensor_against_finite_differences():
    """G = (k^2 I + grad grad) e^{ikr}/r, differentiated numerically."""
    k = 0.7
    h = 1e-4
    for r in ([1.0, 0.0, 0.0], [1.0, 2.0, -1.5], [0.3, -0.4, 2.2]):
        r = np.array(r)
        for mu, nu in [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]:
            e_mu = np.eye(3)[mu] * h
            e_nu = np.eye(3)[nu] * h
            d2 = (_scalar_green(r + e_mu + e_nu, k) - _scalar_green(r + e_mu - e_nu, k)
                  - _scalar_green(r - e_mu + e_nu, k) + _scalar_green(r - e_mu - e_nu, k)) / (4 * h * h)
            ref = d2 + (k * k * _scalar_green(r, k) if mu == nu else 0)
            got = O.green(r, k, mu, nu)
            assert abs(got - ref) <= 1e-6 * max(1.0, abs(ref)), (r, mu, nu, got, ref)
    # static limit at r = (1,0,0): G_xx = 2, G_yy = -1
    assert abs(O.green([1, 0, 0], 0.0, 0, 0) - 2) < 1e-15
    assert abs(O.green([1, 0, 0], 0.0, 1, 1) + 1) < 1e-15


def test_dmatrix_is_fft_of_clipped_green():
    """DMatrixCpu::createDMatrix == numpy.fft.fftn of the clipped tensor / (-N)."""
    box = np.array([4, 3, 2], np.uint32)
    grid = np.array([8, 8, 6], np.uint32)
    d, k = 1.0, 0.6
    dm = O.dmatrix(box, grid, d, k)
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    gx, gy, gz = (int(v) for v in grid)
    bx, by, bz = (int(v) for v in box)
    for c, (mu, nu) in enumerate(comps):
        g = np.zeros((gx, gy, gz), complex)
        for i in range(1 - bx, bx):
            for j in range(1 - by, by):
                for kk in range(1 - bz, bz):
                    if i or j or kk:
                        g[i % gx, j % gy, kk % gz] = O.green([i * d, j * d, kk * d], k, mu, nu)
        ref = np.fft.fftn(g) / -(gx * gy * gz)
        got = dm[:, :, :, c].transpose(0, 2, 1)  # [Gx, Gz, Gy] -> [Gx, Gy, Gz]
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def _problem(shape, p, d=0.5e-6, lam=6e-6, m=1.5 + 0.1j, pot=False, grid=None):
    pos, box = O.geometry(shape, *p)
    g = np.array([O.fft_fit(2 * int(b), 0, not pot) for b in box], np.uint32) if grid is None else np.array(grid)
    prop, pol1, pol2 = O.beam_pols([0, 0, 1])
    k = 2 * math.pi / lam
    cc, cs = O.couple_constants("ldr", m, d, k * d, prop, pol1)
    return O.Problem(pos, None, box, g, cs, d, k), prop, pol1, cc


def _rand(P, seed=3):
    rng = np.random.default_rng(seed)
    x = np.zeros(3 * P.vec_stride, complex)
    for c in range(3):
        x[c * P.vec_stride: c * P.vec_stride + P.nv] = rng.uniform(-1, 1, P.nv) + 1j * rng.uniform(-1, 1, P.nv)
    return x


@pytest.mark.parametrize("shape,p", [("sphere", (3.0,)), ("box", (5, 4, 3)), ("ellipsoid", (4, 3, 2)),
                                     ("cylinder", (6, 2.5))])
@pytest.mark.parametrize("conj", [False, True])
def test_matvec_against_direct_sum(shape, p, conj):
    P, *_ = _problem(shape, p + (0.0,) * (3 - len(p)))
    x = _rand(P)
    y = P.matvec(x, conj=conj, precision=80)
    y_ref = P.matvec_direct(x, conj=conj)
    assert np.linalg.norm(y - y_ref) <= 1e-14 * np.linalg.norm(y_ref)
    assert np.all(y[P.nv:P.vec_stride] == 0)


def test_matvec_direct_sum_on_reference_geometry_fixture():
    """The 480-dipole 10x8x6 box of DDA/box_10_8_6.geom."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "box_10_8_6.geom")
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    pos = np.array(rows, np.uint32)[:, :3]
    assert pos.shape == (480, 3)
    box = pos.max(axis=0) + 1
    d, k = 1.0, 0.45
    prop, pol1, _ = O.beam_pols([0, 0, 1])
    cc, cs = O.couple_constants("cm", 1.33 + 0.01j, d, k * d, prop, pol1)
    P = O.Problem(pos, None, box, np.array([O.fft_fit(2 * int(b)) for b in box]), cs, d, k)
    x = _rand(P, 5)
    assert np.linalg.norm(P.matvec(x) - P.matvec_direct(x)) <= 1e-14 * np.linalg.norm(x)


def test_matvec_grid_invariance_and_symmetry():
    P1, *_ = _problem("sphere", (3.5, 0, 0))
    P2, *_ = _problem("sphere", (3.5, 0, 0), pot=True)
    P3, *_ = _problem("sphere", (3.5, 0, 0), grid=[30, 18, 16])
    x, y = _rand(P1, 1), _rand(P1, 2)
    a = P1.matvec(x)
    assert np.linalg.norm(P2.matvec(x) - a) <= 1e-14 * np.linalg.norm(a)
    assert np.linalg.norm(P3.matvec(x) - a) <= 1e-14 * np.linalg.norm(a)
    # complex symmetry of A = I - sqrt(a) G sqrt(a): x^T A y = y^T A x
    assert abs(np.dot(x, P1.matvec(y)) - np.dot(y, a)) <= 1e-13 * abs(np.dot(y, a))


def _dense(P):
    n = 3 * P.vec_stride
    idx = [c * P.vec_stride + i for c in range(3) for i in range(P.nv)]
    A = np.zeros((len(idx), len(idx)), complex)
    for j, col in enumerate(idx):
        e = np.zeros(n, complex)
        e[col] = 1
        A[:, j] = P.matvec(e)[idx]
    return A, idx


@pytest.mark.parametrize("kind", ["qmr", "bicgstab", "cgnr", "bicg"])
def test_solvers_against_dense_lu(kind):
    P, prop, pol1, cc = _problem("sphere", (2.5, 0, 0))
    A, idx = _dense(P)
    e = O.einc_plane(P.pos, P.box, P.vec_stride, P.grid_unit, P.wave_num, prop, pol1)
    b = np.zeros_like(e)
    for c in range(3):
        s = slice(c * P.vec_stride, c * P.vec_stride + P.nv)
        b[s] = P.cc_sqrt[0, c] * e[s]
    x_ref = np.linalg.solve(A, b[idx])
    st, pvec, it, hist = P.solve(kind, e, 1e-12, max_iter=3000)
    assert st == O.ST_OK
    p_ref = np.zeros_like(e)
    p_ref[idx] = x_ref
    for c in range(3):
        s = slice(c * P.vec_stride, c * P.vec_stride + P.nv)
        p_ref[s] *= P.cc_sqrt[0, c]
    assert np.linalg.norm(pvec - p_ref) <= 1e-9 * np.linalg.norm(p_ref)
    assert hist[-1] < 1e-24


def test_solver_error_statuses():
    P, prop, pol1, cc = _problem("sphere", (3.0, 0, 0))
    e = O.einc_plane(P.pos, P.box, P.vec_stride, P.grid_unit, P.wave_num, prop, pol1)
    st, _, it, _ = P.solve("qmr", e, 1e-14, max_iter=3)
    assert st == O.ST_MAXITER and it == 3


def test_optical_theorem_and_mie():
    """Cext from sum Im(p.E*) equals Cext from the forward amplitude, and both
    approach the Mie value as the sphere is resolved."""
    lam = 6e-6
    k = 2 * math.pi / lam
    m = 1.5 + 0.1j
    radius = 1e-6
    errs = []
    for nd in (4, 8):
        d = radius / nd
        P, prop, pol1, cc = _problem("sphere", (float(nd), 0, 0), d=d, lam=lam, m=m)
        e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol1)
        st, pvec, it, _ = P.solve("qmr", e, 1e-12, max_iter=2000)
        assert st == 0
        c1 = P.ext_cross_sum(pvec, e)
        c2 = P.ext_cross_forward(pvec, prop, pol1)
        assert abs(c1 - c2) <= 1e-9 * abs(c1)
        # volume-equivalent radius of the dipole set
        a_eq = (3 * P.nv / (4 * math.pi)) ** (1 / 3) * d
        qext, qsca, qback, gsca = O.mie(k * a_eq, m)
        errs.append(abs(c1 - qext * math.pi * a_eq ** 2) / (qext * math.pi * a_eq ** 2))
    assert errs[-1] < 0.05
    # Cext = Cabs + Csca with Csca from the far field on a (theta, phi) grid
    nt, nph = 61, 48
    th = np.linspace(0, math.pi, nt)
    J = P.jones(pvec, None, True, prop, pol1, O.beam_pols(prop)[2], nt, 0.0, th[1] - th[0], nph, 0.0,
                2 * math.pi / nph)
    assert np.all(np.isfinite(J))


def test_mie_limits():
    # Rayleigh limit: Qabs ~ 4 x Im((m^2-1)/(m^2+2)), Qsca ~ 8/3 x^4 |(m^2-1)/(m^2+2)|^2
    m = 1.5 + 0.1j
    x = 1e-3
    a = (m * m - 1) / (m * m + 2)
    qext, qsca, qback, g = O.mie(x, m)
    assert abs((qext - qsca) / (4 * x * a.imag) - 1) < 1e-5
    assert abs(qsca / (8.0 / 3 * x ** 4 * abs(a) ** 2) - 1) < 1e-4
    # no absorption: Qext == Qsca
    qext, qsca, qback, g = O.mie(3.0, 1.33)
    assert abs(qext - qsca) < 1e-12 * qext
    # extinction paradox: Qext -> 2 for large x
    qext, *_ = O.mie(20000.0, 1.33 + 0.01j)
    assert abs(qext - 2) < 0.005
    # Bohren & Huffman's BHMIE sample case (sphere radius 0.525 um, wavelength
    # 0.6328 um, m = 1.55): QEXT = QSCA = 3.10543, QBACK = 2.92534, g = 0.63314
    qext, qsca, qback, g = O.mie(2 * math.pi * 0.525 / 0.6328, 1.55)
    assert abs(qext / 3.10543 - 1) < 2e-5 and abs(qsca / 3.10543 - 1) < 2e-5
    assert abs(qback / 2.92534 - 1) < 2e-5 and abs(g / 0.63314 - 1) < 2e-5


def test_ldr_limits():
    d = 1.0
    prop, pol1, _ = O.beam_pols([0, 0, 1])
    m = 1.7
    cm, _ = O.couple_constants("cm", m, d, 1e-3, prop, pol1)
    ldr, _ = O.couple_constants("ldr", m, d, 1e-3, prop, pol1)
    assert abs(ldr[0, 0] / cm[0, 0] - 1) < 1e-5  # LDR -> CM as kd -> 0
    # radiative reaction: for real m, Im(1/a_LDR) - Im(1/a_CM) = -(2/3) k^3
    kd = 0.4
    cm, _ = O.couple_constants("cm", m, d, kd, prop, pol1)
    ldr, _ = O.couple_constants("ldr", m, d, kd, prop, pol1)
    assert abs((1 / ldr[0, 0]).imag - (1 / cm[0, 0]).imag + 2.0 / 3 * kd ** 3) < 1e-12
    # CM: a / V -> 3/(4 pi) as m -> infinity, and 0 at m = 1
    cmi, _ = O.couple_constants("cm", 1e6, d, kd, prop, pol1)
    assert abs(cmi[0, 0] - 3 / (4 * math.pi)) < 1e-9
    one, _ = O.couple_constants("cm", 1.0, d, kd, prop, pol1)
    assert abs(one[0, 0]) == 0


def test_geometry_counts():
    # the published workload sizes: 40 across -> 33,552 dipoles; 128 across -> 1,099,136
    pos, box = O.geometry("sphere", 20.0)
    assert len(pos) == 33552 and list(box) == [40, 40, 40]
    pos, box = O.geometry("box", 10, 8, 6)
    assert len(pos) == 480
    # sorted by (z, y, x)
    key = pos[:, 2].astype(np.int64) * 10 ** 8 + pos[:, 1] * 10 ** 4 + pos[:, 0]
    assert np.all(np.diff(key) > 0)


def test_oracle_matvec_thread_count_independent():
    """The OpenMP matvec (the multi-core CPU baseline) gives bitwise the serial result."""
    import cases
    from dda_amd import problem as P
    g, params, cc, cs, prop, pol = cases.build("sphere_r7")
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=12)
    O.set_threads(1)
    y1 = orc.matvec(x, precision=64)
    O.set_threads(4)
    y4 = orc.matvec(x, precision=64)
    O.set_threads(1)
    assert np.array_equal(y1, y4)


def _csca(P, pvec, nt=24, nph=48):
    """Integral of |E_far|^2 over the unit sphere: Gauss-Legendre in cos(theta),
    uniform (spectrally accurate) in phi."""
    x, w = np.polynomial.legendre.leggauss(nt)
    tot = 0.0
    for ct, wt in zip(x, w):
        st = math.sqrt(max(0.0, 1 - ct * ct))
        for j in range(nph):
            ph = 2 * math.pi * j / nph
            e = P.calc_field(pvec, (st * math.cos(ph), st * math.sin(ph), ct))
            tot += wt * (2 * math.pi / nph) * float(np.sum(np.abs(e) ** 2))
    return tot


@pytest.mark.parametrize("shape,p,m", [("box", (5, 4, 3), 1.5 + 0.1j), ("sphere", (3.0, 0, 0), 1.5 + 0.1j),
                                       ("box", (5, 4, 3), 1.7)])
def test_energy_balance_and_jones_normalisation(shape, p, m):
    """Energy conservation of the dipole solution pins the observables against
    each other: Cext (the sum of Im(p.E*), already pinned by the optical
    theorem) = Cabs + Csca per polarisation, with Csca the integral of |E_far|^2
    over all directions (a quadrature independent of any of the three), and
    Cabs = 0 for a real m (LDR's radiative reaction). The Jones matrices of a
    grid of angles integrate to Csca(pol 1) + Csca(pol 2)."""
    d, lam = 0.5e-6, 6e-6
    k = 2 * math.pi / lam
    pos, box = O.geometry(shape, *p)
    g = np.array([O.fft_fit(2 * int(b)) for b in box], np.uint32)
    prop, pol1, pol2 = O.beam_pols([0, 0, 1])
    csca = []
    ps = []
    for pol in (pol1, pol2):
        cc, cs = O.couple_constants("ldr", m, d, k * d, prop, pol)
        P = O.Problem(pos, None, box, g, cs, d, k)
        e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol)
        st, pvec, _, _ = P.solve("qmr", e, 1e-12, max_iter=3000)
        assert st == 0
        cext = P.ext_cross_sum(pvec, e)
        cabs = P.abs_cross(pvec, cc)
        cs_ = _csca(P, pvec)
        assert abs(cext - (cabs + cs_)) <= 1e-7 * cext, (cext, cabs, cs_)
        if np.imag(m) == 0:
            assert abs(cabs) <= 1e-9 * cext
        else:
            assert cabs > 0.1 * cext
        csca.append(cs_)
        ps.append(pvec)
    # Jones: composite Simpson in theta (sin weight), periodic trapezoid in phi
    nt, nph = 181, 48
    dt, dp = math.pi / (nt - 1), 2 * math.pi / nph
    J = P.jones(ps[0], ps[1], 0, prop, pol1, pol2, nt, 0.0, dt, nph, 0.0, dp).reshape(nt, nph, 4)
    wt = np.ones(nt)
    wt[1:-1:2], wt[2:-1:2] = 4, 2
    wt *= dt / 3 * np.sin(np.arange(nt) * dt)
    tot = float(np.sum(wt[:, None] * np.sum(np.abs(J) ** 2, axis=2)) * dp)
    assert abs(tot - (csca[0] + csca[1])) <= 1e-6 * tot, (tot, csca)


def test_jones_element_placement_on_a_mirror_symmetric_target():
    """The lattice sphere is mirror-symmetric about the planes through its centre
    that contain the propagation axis and a lattice axis, so in those scattering
    planes (phi = 0 and 90 degrees) the off-diagonal Jones elements (s3 = J[0][1],
    s4 = J[1][0]) vanish and the diagonal ones do not; in the forward direction
    the two diagonal elements are equal and give Cext by the optical theorem:
    Cext = -(4 pi / k) Re J_diag(0) for the polarisation along that element."""
    d, lam, m = 0.5e-6, 6e-6, 1.5 + 0.1j
    k = 2 * math.pi / lam
    pos, box = O.geometry("sphere", 3.0)
    g = np.array([O.fft_fit(2 * int(b)) for b in box], np.uint32)
    prop, pol1, pol2 = O.beam_pols([0, 0, 1])
    ps, cext = [], []
    for pol in (pol1, pol2):
        cc, cs = O.couple_constants("ldr", m, d, k * d, prop, pol)
        P = O.Problem(pos, None, box, g, cs, d, k)
        e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol)
        st, pvec, _, _ = P.solve("qmr", e, 1e-12, max_iter=3000)
        ps.append(pvec)
        cext.append(P.ext_cross_sum(pvec, e))
    nt = 19
    J = P.jones(ps[0], ps[1], 0, prop, pol1, pol2, nt, 0.0, math.pi / (nt - 1), 2, 0.0, math.pi / 2)
    J = J.reshape(nt, 2, 2, 2)  # [theta][phi 0, 90][row][col]
    diag = np.abs(J[:, :, 0, 0]) + np.abs(J[:, :, 1, 1])
    off = np.abs(J[:, :, 0, 1]) + np.abs(J[:, :, 1, 0])
    assert np.all(off <= 1e-9 * diag.max()) and np.all(diag > 1e-3 * diag.max())
    f = J[0, 0]
    assert abs(f[0, 0] - f[1, 1]) <= 1e-9 * abs(f[0, 0])
    # the sphere scatters both polarisations alike: Cext equal, and from the forward amplitude
    assert abs(cext[0] - cext[1]) <= 1e-9 * cext[0]
    assert abs(-4 * math.pi / k * f[0, 0].real - cext[0]) <= 1e-8 * cext[0]


def test_beam_pols_and_plane_wave():
    """Beam.cpp:34-51: (prop, pol1, pol2) is an orthonormal triad for any
    direction (pol1 = y, pol2 = x for +z, the reference's own choice); Beam.cpp:110-118:
    the plane wave has unit amplitude along pol at every dipole and advances in
    phase by k d per grid step along prop, with phase 0 at the box centre."""
    rng = np.random.default_rng(9)
    for v in [(0, 0, 1), (0, 0, -1), (1, 0, 0)] + [tuple(rng.normal(size=3)) for _ in range(5)]:
        prop, p1, p2 = O.beam_pols(v)
        M = np.array([prop, p1, p2])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-14)
        assert np.allclose(prop, np.asarray(v) / np.linalg.norm(v), atol=1e-14)
        # pol1 is horizontal (no z part) and pol2 = pol1 x prop (ADDA's frame)
        assert abs(p1[2]) <= 1e-14 and np.allclose(p2, np.cross(p1, prop), atol=1e-14)
    prop, p1, p2 = O.beam_pols([0, 0, 1])
    assert np.allclose(p1, [0, 1, 0]) and np.allclose(p2, [1, 0, 0])
    _, q1, q2 = O.beam_pols([0, 0, -1])
    assert np.allclose(q1, [0, 1, 0]) and np.allclose(q2, [-1, 0, 0])
    pos, box = O.geometry("box", 4, 3, 5)
    d, k = 0.5e-6, 2 * math.pi / 6e-6
    vs = (len(pos) + 15) // 16 * 16
    e = O.einc_plane(pos, box, vs, d, k, prop, p1)
    E = np.stack([e[c * vs: c * vs + len(pos)] for c in range(3)], axis=1)
    assert np.allclose(np.abs(E[:, 1]), 1, atol=1e-15) and np.all(E[:, 0] == 0) and np.all(E[:, 2] == 0)
    z = pos[:, 2].astype(float) - (box[2] - 1) / 2.0
    assert np.allclose(E[:, 1], np.exp(1j * k * d * z), atol=1e-14)


# The LDR constants of FPConst.hpp:48-50 (doi:10.1086/172396), used only to
# plant variants below; the oracle's own values are not read from here.
_LDR_B = (1.8915316, -0.1648469, 1.7700004)


def _sphere_cabs_error(nd, m, mkd, prop_dir, planted=None, kind="ldr"):
    """Mean relative error of Cabs (both polarisations) of the lattice sphere of
    radius nd dipoles against Mie theory for the volume-equivalent sphere."""
    d = 1.0
    kd = mkd / abs(m)
    k = kd / d
    pos, box = O.geometry("sphere", float(nd))
    grid = np.array([O.fft_fit(2 * int(b)) for b in box], np.uint32)
    prop, pol1, pol2 = O.beam_pols(np.asarray(prop_dir, float) / np.linalg.norm(prop_dir))
    a_eq = (3 * len(pos) / (4 * math.pi)) ** (1 / 3) * d
    qext, qsca, _, _ = O.mie(k * a_eq, m)
    cabs_mie = (qext - qsca) * math.pi * a_eq ** 2
    dmat, cabs = None, []
    for pol in (pol1, pol2):
        cc, cs = O.couple_constants(kind, m, d, kd, prop, pol, planted=planted)
        P = O.Problem(pos, None, box, grid, cs, d, k, dmat=dmat)
        dmat = P.dmat
        e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol)
        st, pv, _, _ = P.solve("qmr", e, 1e-9, max_iter=5000)
        assert st == O.ST_OK
        cabs.append(P.abs_cross(pv, cc))
    return abs(np.mean(cabs) / cabs_mie - 1)


def test_ldr_finite_kd_terms_against_mie_at_oblique_incidence():
    """Pins the O((kd)^2) part of the LDR polarisability (b1, b2, b3 and S =
    sum prop_i^2 pol_i^2, PolarizabilityDescription.cpp:85-110) by physics: the
    LDR makes the dipole lattice propagate light with the medium's own index,
    so the absorption of a lattice sphere follows Mie theory closely. Along
    (1,1,1)/sqrt(3), S = 1/3 for every polarisation. The true constants meet a
    0.7% tolerance (measured 0.26%); each planted mistake -- S dropped, b3, b2 or
    b1 with the wrong sign -- misses it (measured 1.4% to 4.4%)."""
    O.set_threads(8)
    case = (12, 1.5 + 0.1j, 0.5, (1, 1, 1))
    tol = 7e-3
    assert _sphere_cabs_error(*case) < tol
    b1, b2, b3 = _LDR_B
    for planted in ((b1, b2, b3, 0.0), (b1, b2, -b3, 1.0), (b1, -b2, b3, 1.0), (-b1, b2, b3, 1.0)):
        assert _sphere_cabs_error(*case, planted=planted) > tol, planted
    # Clausius-Mossotti (no lattice correction at all) misses it too
    assert _sphere_cabs_error(*case, kind="cm") > tol
    O.set_threads(1)


def test_ldr_avgpol_is_the_polarisation_mean():
    """1/alpha is linear in S, and LDR-avgpol's S = (1 - sum prop_i^4) / 2 is the
    mean of S = sum prop_i^2 pol_i^2 over two orthogonal polarisations across
    prop (sum over pol of pol_i^2 = 1 - prop_i^2), so 1/alpha_avgpol is the mean
    of 1/alpha over polarisations 1 and 2, for any direction."""
    rng = np.random.default_rng(4)
    m = [1.33 + 0.01j, 2.0 + 0.4j]
    for _ in range(4):
        prop, p1, p2 = O.beam_pols(rng.normal(size=3))
        a1, _ = O.couple_constants("ldr", m, 1.0, 0.6, prop, p1)
        a2, _ = O.couple_constants("ldr", m, 1.0, 0.6, prop, p2)
        av, _ = O.couple_constants("ldr-avgpol", m, 1.0, 0.6, prop, p1)
        np.testing.assert_allclose(1 / av, (1 / a1 + 1 / a2) / 2, rtol=1e-13)
        # and it does depend on the direction's S (not a constant)
        assert np.max(np.abs(1 / a1 - 1 / a2)) > 1e-3 * np.max(np.abs(1 / a1))


def pinwheel_target():
    import cases
    pos = cases.pinwheel_positions()
    return pos, pos.max(axis=0) + 1


def test_symmetric_jones_equals_two_solves():
    """FarFieldCalc.cpp:163-168: for a target symmetric under the quarter turn
    about the propagation axis that takes polarisation 1 to polarisation 2, the
    field of polarisation 2 is that of polarisation 1 with the axes (X, Y) =
    (pol1, -pol2). The chiral pinwheel along z is such a target (LDR gives both
    polarisations the same polarisability there, S = 0), so the Jones matrices
    from one solve with symmetric = 1 equal those from both solves to the
    solver's tolerance. Because the target has no mirror plane, dropping the
    minus on pol2 (the mirror image) or swapping the two axes -- mistakes
    planted in the oracle -- gives different matrices."""
    d, lam, m = 0.5e-6, 3e-6, 1.5 + 0.1j
    k = 2 * math.pi / lam
    pos, box = pinwheel_target()
    g = np.array([O.fft_fit(2 * int(b)) for b in box], np.uint32)
    prop, pol1, pol2 = O.beam_pols([0, 0, 1])
    ps = []
    for pol in (pol1, pol2):
        cc, cs = O.couple_constants("ldr", m, d, k * d, prop, pol)
        P = O.Problem(pos, None, box, g, cs, d, k)
        e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol)
        st, pvec, _, _ = P.solve("qmr", e, 1e-12, max_iter=3000)
        assert st == O.ST_OK
        ps.append(pvec)
    args = (prop, pol1, pol2, 13, 0.0, math.pi / 12, 12, 0.0, math.pi / 6)
    J2 = P.jones(ps[0], ps[1], 0, *args)
    J1 = P.jones(ps[0], None, 1, *args)
    scale = np.max(np.abs(J2))
    assert np.max(np.abs(J1 - J2)) <= 1e-9 * scale
    for planted in (2, 3):
        Jbad = P.jones(ps[0], None, planted, *args)
        assert np.max(np.abs(Jbad - J2)) > 1e-2 * scale, planted


def test_solver_breakdown_status():
    """A right-hand side with b^T b = 0 (b != 0: E_x = 1 at one dipole, i at the
    next) started from x0 = 0 makes BiCG-CS's rho = r^T r zero at its first
    iteration, which the reference stops on (BicgCs.cpp:68): status BREAKDOWN
    after 0 iterations, with x still 0."""
    P, prop, pol1, cc = _problem("sphere", (3.0, 0, 0))
    e = np.zeros(3 * P.vec_stride, complex)
    e[0], e[1] = 1.0, 1.0j
    st, pv, it, _ = P.solve("bicg", e, 1e-8, max_iter=100, start=np.zeros_like(e))
    assert st == O.ST_BREAKDOWN and it == 0
    assert np.all(pv == 0)


def _stagnation_case():
    d, lam, m = 0.5e-6, 6e-6, 3.0 + 0.01j
    k = 2 * math.pi / lam
    pos, box = O.geometry("sphere", 8.0)
    g = np.array([O.fft_fit(2 * int(b)) for b in box], np.uint32)
    prop, pol1, _ = O.beam_pols([0, 0, 1])
    cc, cs = O.couple_constants("ldr", m, d, k * d, prop, pol1)
    P = O.Problem(pos, None, box, g, cs, d, k)
    e = O.einc_plane(P.pos, P.box, P.vec_stride, d, k, prop, pol1)
    return P, e


@pytest.mark.parametrize("kind,mri,it_expected", [("bicg", 1, 6), ("bicg", 2, 12), ("qmr", 1, 27)])
def test_solver_stagnation_status(kind, mri, it_expected):
    """IterativeSolverBase.cpp:56-70: the solve stops with "Too many iterations w/o
    increase" once more than maxResIncrease iterations in a row failed to lower
    the smallest residual seen. On a high-index sphere (m = 3 + 0.01i, whose
    residuals are erratic) the stop comes at the first iteration where the
    residual history returned by the solve has mri + 1 such iterations in a row,
    well before convergence."""
    P, e = _stagnation_case()
    st, _, it, hist = P.solve(kind, e, 1e-10, max_iter=3000, max_res_increase=mri)
    assert st == O.ST_STAGNATION and it == it_expected
    best, run = hist[0], 0
    for i, h in enumerate(hist[1:], 1):
        run = 0 if h <= best else run + 1
        best = min(best, h)
        if run > mri:
            break
    assert i == it
    # the same solve with the default limit goes on (no stagnation that early)
    st2, _, it2, _ = P.solve(kind, e, 1e-10, max_iter=it + 5)
    assert st2 != O.ST_STAGNATION


@pytest.mark.parametrize("name", ["two_materials", "box_10_8_6", "cylinder_9_3"])
def test_pocketfft_matvec_equals_oracle_matvec(name):
    """The CPU baseline's MatVecCpu restatement over scipy.fft (pocketfft, the
    stand-in for the reference's FFTW) gives the oracle's matvec to rounding,
    both conj flags, serial and threaded."""
    import cases
    from dda_amd import problem as P
    g, params, cc, cs, prop, pol = cases.build(name)
    orc = cases.oracle_problem(g, params, cs, dmat_precision=64)
    x = P.random_dipole_vector(params, seed=5)
    for conj in (False, True):
        y0 = orc.matvec(x, conj=conj)
        for w in (1, 3):
            y1 = orc.matvec_pocketfft(x, conj=conj, workers=w)
            assert np.linalg.norm(y1 - y0) <= 1e-14 * np.linalg.norm(y0)
            assert np.all(y1.reshape(3, -1)[:, params.nv:] == 0)
