"""Pins for the CPU oracle (oracle/pm_oracle.c).

Each function of the oracle is checked against something other than its own
formula: the reference's golden vectors and its vendored kissfft, closed
forms, moment identities, Gauss's law for the S1/S2 reference forces, the
continuum Fourier transforms of the S1/S2 clouds, exact discrete-Laplacian
inversion, brute-force pair sums, and the momentum identity of a symmetric
PM scheme.
"""
import json
import os

import numpy as np
import pytest
from scipy import integrate

GOLD = os.path.join(os.path.dirname(__file__), "golden")
GC = 1.0 / (4.0 * np.pi)


def _rng(seed=0):
    return np.random.default_rng(seed)


# --------------------------------------------------------------------------
# deposit (source/pmMethod.cpp:200-277)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_deposit_conserves_mass(oracle, scheme):
    r = _rng(1)
    pos = r.uniform(3, 13, (500, 3))
    m = r.uniform(0.5, 2.0, 500)
    rho = oracle.deposit(pos, m, (16, 16, 16), scheme)
    assert abs(rho.sum() - m.astype(np.float32).astype(np.float64).sum()) < 1e-9 * m.sum()


@pytest.mark.parametrize("scheme", [1, 2, 3])
def test_deposit_first_moment_is_position(oracle, scheme):
    # sum_w w*(anchor+offset) = x for CIC and (floor-anchored) TSC
    r = _rng(2)
    pos = r.uniform(4, 12, (200, 3)).astype(np.float32)
    m = np.ones(200, np.float32)
    rho = oracle.deposit(pos, m, (16, 16, 16), scheme)
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    for axis, coord in ((0, x), (1, y), (2, z)):
        assert abs((rho * coord).sum() - pos[:, axis].astype(np.float64).sum()) < 1e-9 * 200 * 16


@pytest.mark.parametrize("delta", [0.0, 0.25, 0.5, 0.9, 0.999])
def test_tsc_weights_closed_form(oracle, delta):
    # floor-anchored TSC (App. Q1): weights 1/2(1/2-d)^2, 3/4-d^2, 1/2(1/2+d)^2
    # sum 1, mean d, variance 1/4 -- a swapped or mis-signed weight fails
    pos = np.array([[5 + delta, 5.0, 5.0]], np.float32)
    d = float(np.float32(5 + delta) - 5)
    rho = oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 2)
    w = rho[5, 5, 4:7] / (0.75 ** 2)  # y, z at delta 0 carry 3/4 each
    assert np.allclose(w.sum(), 1.0, atol=1e-12)
    assert np.isclose((w * np.array([-1, 0, 1])).sum(), d, atol=1e-12)
    assert np.isclose((w * (np.array([-1, 0, 1]) - d) ** 2).sum(), 0.25, atol=1e-12)
    if delta == 0.25:
        assert np.allclose(w, [0.03125, 0.6875, 0.28125], atol=1e-12)


@pytest.mark.parametrize("delta", [0.0, 0.25, 0.49, 0.5, 0.75, 0.999])
def test_centred_tsc_weights_closed_form(oracle, delta):
    # textbook TSC (an extension, scheme 3): the quadratics about the nearest
    # grid point, d = x - round(x) in [-1/2, 1/2]: weights never negative,
    # sum 1, mean at the particle, variance d^2 + 1/4 - d^2 = 1/4 about it
    pos = np.array([[5 + delta, 5.0, 5.0]], np.float32)
    x = float(np.float32(5 + delta))
    c = int(np.floor(x + 0.5))
    d = x - c
    rho = oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 3)
    w = rho[5, 5, c - 1:c + 2] / (0.75 ** 2)
    assert rho[5, 5].sum() / 0.75 ** 2 == pytest.approx(1.0, abs=1e-12)  # nothing outside c-1..c+1
    assert np.all(w >= 0.0)
    assert np.allclose(w, [0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2], atol=1e-12)
    assert np.isclose((w * (np.arange(c - 1, c + 2) - x)).sum(), 0.0, atol=1e-12)
    assert np.isclose((w * (np.arange(c - 1, c + 2) - x) ** 2).sum(), 0.25, atol=1e-12)


def test_centred_tsc_equals_floor_tsc_below_half(oracle):
    # the two anchors agree wherever floor(x) == round(x): fractional parts
    # below 1/2 on every axis give identical densities, above 1/2 they differ
    r = _rng(9)
    base = r.integers(4, 12, (300, 3)).astype(np.float64)
    lo = (base + r.uniform(0.0, 0.45, (300, 3))).astype(np.float32)
    hi = (base + r.uniform(0.55, 0.95, (300, 3))).astype(np.float32)
    m = r.uniform(0.5, 1.5, 300).astype(np.float32)
    assert np.array_equal(oracle.deposit(lo, m, (16, 16, 16), 3), oracle.deposit(lo, m, (16, 16, 16), 2))
    a, b = oracle.deposit(hi, m, (16, 16, 16), 3), oracle.deposit(hi, m, (16, 16, 16), 2)
    assert np.abs(a - b).max() > 0.1 * np.abs(a).max()
    assert b.min() < 0.0 <= a.min()  # floor-anchored centre weights go negative above sqrt(3)/2


def test_cic_weights_closed_form(oracle):
    pos = np.array([[5.3, 6.6, 7.0]], np.float32)
    rho = oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 1)
    dx, dy = float(np.float32(5.3)) - 5, float(np.float32(6.6)) - 6
    assert np.isclose(rho[7, 6, 5], (1 - dx) * (1 - dy), atol=1e-12)
    assert np.isclose(rho[7, 7, 6], dx * dy, atol=1e-12)
    assert rho[8].sum() == 0.0  # dz = 0


def test_ngp_rounds(oracle):
    pos = np.array([[5.49, 5.5, 5.51]], np.float32)
    rho = oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 0)
    assert rho[6, 6, 5] == 1.0  # round half away from zero (std::round)


def test_deposit_unwrapped_rejects_out_of_grid(oracle):
    pos = np.array([[0.2, 5, 5]], np.float32)  # TSC reaches cell -1: UB in the reference
    with pytest.raises(ValueError):
        oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 2, wrap=False)
    rho = oracle.deposit(pos, np.ones(1, np.float32), (16, 16, 16), 2, wrap=True)
    assert rho[5, 5, 15] > 0


# --------------------------------------------------------------------------
# FFT (FFTWAdapter / KissFFTAdapter)
# --------------------------------------------------------------------------

def test_fft_golden_roundtrip_2x2x2(oracle):
    # the reference's only FFT test: test/fftAdaptersTest.cpp:6-22
    with open(os.path.join(GOLD, "fft_roundtrip_2x2x2.json")) as f:
        g = json.load(f)
    a = np.array(g["input"], np.complex128).reshape(2, 2, 2)
    f = oracle.fft3d(a)
    assert np.allclose(f.ravel(), np.array(g["forward_re"]) + 1j * np.array(g["forward_im"]), atol=1e-12)
    b = oracle.fft3d(f, inverse=True) / a.size
    assert np.max(np.abs(b - a)) <= g["tolerance"]


def test_fft_matches_kissfft_fixture(oracle):
    # forward spectrum of the reference's own FFT backend (kissfft, built from
    # the reference tree by oracle/Makefile) committed as a fixture
    d = np.load(os.path.join(GOLD, "kissfft_8x16x32.npz"))
    f = oracle.fft3d(d["input"])
    assert np.max(np.abs(f - d["forward"])) < 2e-4 * np.max(np.abs(d["forward"]))
    b = oracle.fft3d(d["forward"].astype(np.complex128), inverse=True) / d["input"].size
    assert np.max(np.abs(b - d["inverse_roundtrip"])) < 1e-5


@pytest.mark.skipif("not __import__('oracle.oracle', fromlist=['x']).kissfft_available()")
@pytest.mark.parametrize("shape", [(16, 8, 32), (30, 60, 60), (15, 34, 68)])
def test_fft_matches_live_kissfft(oracle, shape):
    # the reference's own mixed-radix FFT backend, built from its sources
    a = (_rng(3).standard_normal(shape) + 1j * _rng(4).standard_normal(shape)).astype(np.complex64)
    k = oracle.kiss_fftnd(a)
    f = oracle.fft3d(a)
    assert np.max(np.abs(f - k)) < 1e-5 * np.max(np.abs(k))
    kb = oracle.kiss_fftnd(k, inverse=True)
    assert np.max(np.abs(kb - a)) < 1e-5


@pytest.mark.parametrize("shape", [(6, 10, 15), (30, 60, 60), (7, 16, 9), (34, 68, 68)])
def test_fft_any_length_matches_numpy(oracle, shape):
    # lengths that are not powers of two take the direct DFT (the reference's
    # pmAccuracy demo sweeps 2Ng x 2Ng x Ng grids, Ng = 30 .. 70, through FFTW,
    # source/demos.cpp:381-408); numpy's pocketfft is an independent FFT
    a = _rng(8).standard_normal(shape) + 1j * _rng(9).standard_normal(shape)
    f = oracle.fft3d(a)
    assert np.max(np.abs(f - np.fft.fftn(a))) <= 1e-10 * np.abs(f).max()
    b = oracle.fft3d(a, inverse=True)
    assert np.max(np.abs(b - np.fft.ifftn(a) * a.size)) <= 1e-10 * np.abs(b).max()


def test_fft_matches_numpy(oracle):
    a = _rng(5).standard_normal((8, 16, 32)) + 1j * _rng(6).standard_normal((8, 16, 32))
    assert np.allclose(oracle.fft3d(a), np.fft.fftn(a), atol=1e-9)
    assert np.allclose(oracle.fft3d(a, inverse=True), np.fft.ifftn(a) * a.size, atol=1e-9)


# --------------------------------------------------------------------------
# Green's functions (source/greensFunctions.cpp)
# --------------------------------------------------------------------------

def _solve(oracle, rho, G):
    ph = oracle.fft3d(oracle.fft3d(rho) * G, inverse=True) / rho.size
    return ph.real


def test_discrete_laplacian_inverts_7point_laplacian(oracle):
    n = 16
    rho = _rng(7).standard_normal((n, n, n))
    G = oracle.green_grid(0, 2, 0, 0.0, (n, n, n))
    phi = _solve(oracle, rho, G)
    lap = -6 * phi
    for ax in range(3):
        lap += np.roll(phi, 1, ax) + np.roll(phi, -1, ax)
    assert np.allclose(lap, rho - rho.mean(), atol=1e-10)


def test_poor_man_solves_continuum_poisson_for_a_plane_wave(oracle):
    n = 32
    x = np.arange(n)
    for m in (1, 3, n - 2):
        rho = np.cos(2 * np.pi * m * x / n)[None, None, :] * np.ones((n, n, 1))
        G = oracle.green_grid(3, 2, 0, 0.0, (n, n, n))
        phi = _solve(oracle, rho, G)
        k = 2 * np.pi * min(m, n - m) / n
        assert np.allclose(phi, -rho / k ** 2, atol=1e-10)


@pytest.mark.parametrize("scheme,lim", [(2, 400), (1, 400), (0, 0)])
def test_alias_sum_closed_forms(oracle, scheme, lim):
    for k in ([0.3, 1.1, 2.9], [6.0, 0.01, 3.14], [1.0, 2.0, 5.5]):
        closed = oracle.alias_sum_closed(scheme, k)
        if scheme == 0:
            assert closed == 1.0
            continue
        brute = oracle.alias_sum_bruteforce(scheme, k, lim)
        assert abs(closed - brute) < 1e-7 * abs(brute)


def _ft_radial(rho_r, R, k):
    # 3-D Fourier transform of a radial density of support R, normalised to mass 1
    mass = integrate.quad(lambda r: 4 * np.pi * r * r * rho_r(r), 0, R, epsabs=1e-13)[0]
    val = integrate.quad(lambda r: 4 * np.pi * r * r * rho_r(r) * np.sinc(k * r / np.pi), 0, R,
                         epsabs=1e-13, limit=200)[0]
    return val / mass


@pytest.mark.parametrize("k", [0.05, 0.4, 1.3, 3.0, 7.5])
def test_s1_s2_fourier_are_cloud_transforms(oracle, k):
    a = 3.0
    # S1: uniform sphere of diameter a; S2: density falling linearly to zero at a/2
    s1 = _ft_radial(lambda r: 1.0, a / 2, k)
    s2 = _ft_radial(lambda r: a / 2 - r, a / 2, k)
    assert abs(oracle.s_fourier(0, k, a) - s1) < 1e-9
    assert abs(oracle.s_fourier(1, k, a) - s2) < 1e-9


def test_s_fourier_series_branch_is_continuous(oracle):
    a = 2.0
    for cs in (0, 1):
        lo = oracle.s_fourier(cs, 0.5 - 1e-12, a)  # u = ka/2 just below the series switch
        hi = oracle.s_fourier(cs, 0.5 + 1e-12, a)
        assert abs(lo - hi) < 1e-12
        assert abs(oracle.s_fourier(cs, 1e-6, a) - 1.0) < 1e-12


def _point_mass_samples(oracle, scheme, fds, G, n, a, rng, reps, rmin=1.0):
    """PM force of a unit point mass at seeded positions on targets at seeded
    directions and distances in [rmin, 2a): radial force, full force vector,
    separation vector and distance per target."""
    rad, acc_all, dvec, dist = [], [], [], []
    for _ in range(reps):
        c = 16.0 + rng.random(3)
        dirs = rng.standard_normal((60, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        r = rng.uniform(rmin, 2 * a, 60)
        pos = np.vstack([c, c + dirs * r[:, None]]).astype(np.float32)
        m = np.zeros(len(pos), np.float32)
        m[0] = 1.0
        acc = oracle.pm_accel(pos, m, (n, n, n), scheme, fds, G)
        d = pos[1:].astype(np.float64) - pos[0]
        dd = np.linalg.norm(d, axis=1)
        rad.append(-(acc[1:] * d).sum(1) / dd)
        acc_all.append(acc[1:])
        dvec.append(d)
        dist.append(dd)
    return tuple(map(np.concatenate, (rad, acc_all, dvec, dist)))


# (scheme, fds, bias, rms, band): NGP's force scatters over the source's cell
# by tens of per cent, so its averages carry looser bounds
_OPT_CASES = [(1, 0, 0.015, 0.06, 0.03), (2, 0, 0.015, 0.06, 0.03), (1, 1, 0.015, 0.06, 0.03),
              (2, 1, 0.015, 0.06, 0.03), (0, 0, 0.09, 0.22, 0.03), (0, 1, 0.09, 0.22, 0.03),
              (3, 0, 0.015, 0.06, 0.03), (3, 1, 0.015, 0.06, 0.03)]


@pytest.mark.parametrize("scheme,fds,bias,rms,band", _OPT_CASES)
def test_optimal_green_reproduces_reference_force(oracle, scheme, fds, bias, rms, band):
    # The reference's own check (script/pm_optimal.py): the PM force of a point
    # mass with the optimal G follows the S1/S2 reference force R(r).  The
    # influence function is optimal on average over source and target
    # positions inside a cell, so the check averages over seeded positions.
    # Periodic images enter at leading order as the background term +m r/(3N^3).
    n, a = 32, 4.0
    rng = _rng(12)
    for gf, cs in ((1, 0), (2, 1)):
        G = oracle.green_grid(gf, scheme, fds, a, (n, n, n))
        rad, _, _, dist = _point_mass_samples(oracle, scheme, fds, G, n, a, rng, 8)
        ref = np.array([oracle.reference_force(cs, float(x), a) for x in dist]) - dist / (3 * n ** 3)
        scale = np.sqrt(np.mean(ref ** 2))
        assert abs(np.mean(rad - ref)) / scale < bias, (scheme, fds, gf)
        assert np.sqrt(np.mean((rad - ref) ** 2)) / scale < rms, (scheme, fds, gf)
        for lo in (4.0, 6.0) if scheme == 0 else (2.0, 4.0, 6.0):
            sel = (dist >= lo) & (dist < lo + 2.0)
            assert abs(np.mean(rad[sel] / ref[sel]) - 1.0) < band, (scheme, fds, gf, lo)


# (scheme, fds) with the optimal G built for a mismatched (scheme', fds'):
# a G that assumes another gradient stencil or another assignment window
_MISMATCH = [((2, 1), (2, 0)), ((2, 0), (2, 1)), ((1, 1), (1, 0)), ((0, 0), (1, 0)), ((0, 1), (1, 1)),
             ((1, 0), (0, 0)), ((1, 0), (2, 0)), ((2, 1), (1, 1))]


def test_centred_tsc_follows_the_reference_force_closer(oracle):
    # the optimal G is built for the TSC window sinc^3 about the nearest grid
    # point; the textbook (centred) TSC has exactly that window, the floor
    # anchor does not, so on the same samples the centred scheme's PM force
    # is closer to the S2 reference force (for both difference stencils)
    n, a = 32, 4.0
    for fds in (0, 1):
        errs = []
        for scheme in (2, 3):
            G = oracle.green_grid(2, scheme, fds, a, (n, n, n))
            _, acc, d, dist = _point_mass_samples(oracle, scheme, fds, G, n, a, _rng(21), 12, rmin=0.5)
            R = np.array([oracle.reference_force(1, float(x), a) for x in dist]) - dist / (3 * n ** 3)
            errs.append(np.sqrt(((acc + R[:, None] * d / dist[:, None]) ** 2).sum() / (R ** 2).sum()))
        assert errs[1] < 0.8 * errs[0], (fds, errs)


@pytest.mark.parametrize("used,built", _MISMATCH)
def test_optimal_green_is_optimal_for_its_scheme_and_stencil(oracle, used, built):
    # GreenOptimal (source/greensFunctions.cpp:122-189) minimises the mean
    # square deviation of the PM force from the reference force over source
    # and target positions.  The G built for the scheme and stencil in use
    # must therefore beat, on the same seeded samples, a G built for another
    # stencil D(k) or another assignment window U(k): a mistake in D4, in the
    # NGP window or in the alias sum of either makes the mismatched G win.
    n, a = 32, 4.0
    scheme, fds = used

    def err(G):
        _, acc, d, dist = _point_mass_samples(oracle, scheme, fds, G, n, a, _rng(21), 12, rmin=0.5)
        R = np.array([oracle.reference_force(1, float(x), a) for x in dist]) - dist / (3 * n ** 3)
        want = -R[:, None] * d / dist[:, None]
        return np.sqrt(((acc - want) ** 2).sum() / (R ** 2).sum())
    good = err(oracle.green_grid(2, scheme, fds, a, (n, n, n)))
    other = err(oracle.green_grid(2, built[0], built[1], a, (n, n, n)))
    assert good < 0.9 * other, (used, built, good, other)


@pytest.mark.parametrize("fds", [0, 1])
def test_fd_d_is_the_field_stencil_on_plane_waves(oracle, fds):
    # D(k) inside GreenOptimal (D2Fourier / D4Fourier, source/greensFunctions.cpp:
    # 70-89) must be the symbol of the field stencil (source/pmMethod.cpp:355-365):
    # for phi = cos(k x) the stencil gives E = -grad phi = d(k) sin(k x).
    n = 32
    x = np.arange(n, dtype=np.float64)
    for m in (1, 3, 7, 12, 16):
        k = 2 * np.pi * m / n
        d = oracle.fd_d(fds, k)
        for ax in range(3):
            shape = [1, 1, 1]
            shape[2 - ax] = n
            phi = np.broadcast_to(np.cos(k * x).reshape(shape), (n, n, n)).copy()
            E = oracle.field(phi, fds)
            want = np.broadcast_to((d * np.sin(k * x)).reshape(shape), (n, n, n))
            assert np.allclose(E[..., ax], want, atol=1e-12), (fds, m, ax)
            others = [c for c in range(3) if c != ax]
            assert np.allclose(E[..., others], 0.0, atol=1e-12)
    # the symbols themselves: 2-point sin k; 4-point (8 sin k - sin 2k)/6
    assert abs(oracle.fd_d(0, 0.7) - np.sin(0.7)) < 1e-15
    assert abs(oracle.fd_d(1, 0.7) - (8 * np.sin(0.7) - np.sin(1.4)) / 6) < 1e-15


# --------------------------------------------------------------------------
# field, interpolation, full PM step
# --------------------------------------------------------------------------

def test_two_point_gradient_exact_on_quadratics(oracle):
    n = 16
    x = np.arange(n, dtype=np.float64)
    phi = np.broadcast_to((x ** 2)[None, None, :], (n, n, n)).copy()
    E = oracle.field(phi, 0)
    assert np.allclose(E[..., 2:-2, 0], -2 * x[2:-2])
    assert np.allclose(E[..., 1:], 0)


def test_four_point_gradient_exact_on_quartics(oracle):
    n = 16
    y = np.arange(n, dtype=np.float64)
    phi = np.broadcast_to((y ** 4 - 3 * y ** 3)[None, :, None], (n, n, n)).copy()
    E = oracle.field(phi, 1)
    assert np.allclose(E[:, 2:-2, :, 1], -(4 * y[2:-2] ** 3 - 9 * y[2:-2] ** 2)[:, None])


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_interpolation_of_constant_and_linear_fields(oracle, scheme):
    n = 16
    pos = _rng(8).uniform(4, 11, (100, 3)).astype(np.float32)
    E = np.zeros((n, n, n, 3))
    E[..., :] = [1.5, -2.0, 0.25]
    acc = oracle.interpolate(pos, E, scheme)
    assert np.allclose(acc, [1.5, -2.0, 0.25], atol=1e-12)
    if scheme > 0:
        E = np.zeros((n, n, n, 3))
        E[..., 0] = np.arange(n)[None, None, :]
        acc = oracle.interpolate(pos, E, scheme)
        assert np.allclose(acc[:, 0], pos[:, 0].astype(np.float64), atol=1e-9)


@pytest.mark.parametrize("scheme,fds,gf", [(0, 0, 0), (1, 0, 0), (2, 1, 0), (1, 1, 3), (2, 0, 2)])
def test_pm_conserves_momentum(oracle, scheme, fds, gf):
    # sum_i m_i a_i = rho^T (D G) rho = 0 for symmetric assignment, antisymmetric D, even G
    n = 32
    r = _rng(9)
    pos = r.uniform(0, n, (400, 3)).astype(np.float32)
    m = r.uniform(0.5, 1.5, 400).astype(np.float32)
    G = oracle.green_grid(gf, scheme, fds, 3.0, (n, n, n))
    acc = oracle.pm_accel(pos, m, (n, n, n), scheme, fds, G)
    p = (m[:, None] * acc).sum(0)
    assert np.max(np.abs(p)) < 1e-10 * np.sum(np.abs(m[:, None] * acc))


def test_pm_point_mass_far_field(oracle):
    n = 64
    G = oracle.green_grid(0, 1, 0, 0.0, (n, n, n))
    c = np.array([32.0, 32.0, 32.0])
    r = np.array([6.0, 8.0, 10.0])
    pos = np.vstack([c, c + np.stack([r, 0 * r, 0 * r], 1)]).astype(np.float32)
    m = np.array([1.0, 0, 0, 0], np.float32)
    acc = oracle.pm_accel(pos, m, (n, n, n), 1, 0, G)
    newton = -GC / r ** 2
    assert np.allclose(acc[1:, 0], newton, rtol=0.06)


def test_external_field_matches_sph_rad_decr(oracle):
    # sphRadDecrField (source/externalFields.cpp:4-16) beyond R is -GM/r^2
    n, H, DT = 16, 0.5, 0.1
    pos = np.array([[12.0, 8.0, 8.0]], np.float32)  # original (6, 4, 4); centre (4,4,4)
    E = np.zeros((n, n, n, 3))
    prm = [4.0, 4.0, 4.0, 1.0, 2.0, 0.7, H, DT]
    acc = oracle.interpolate(pos, E, 2, ext_prm=prm)
    g = -0.7 * 2.0 / 2.0 ** 2
    assert np.allclose(acc[0], [DT * DT * g / H, 0, 0], atol=1e-12)


# --------------------------------------------------------------------------
# P3M short range (source/p3mMethod.cpp)
# --------------------------------------------------------------------------

def _two_sphere_force(cs, r, a):
    """Force between two unit-mass S1/S2 clouds at separation r (G'=1/4pi),
    by direct integration of cloud 1's field over cloud 2."""
    R = a / 2
    if cs == 0:
        dens = lambda s: np.where(s < R, 1.0, 0.0)  # noqa: E731
        menc = lambda s: min(s, R) ** 3 / R ** 3  # noqa: E731
    else:
        dens = lambda s: np.where(s < R, R - s, 0.0)  # noqa: E731
        norm = 4 * np.pi * (R ** 4 / 3 - R ** 4 / 4)

        def menc(s):
            s = min(s, R)
            return 4 * np.pi * (R * s ** 3 / 3 - s ** 4 / 4) / norm
    norm2 = integrate.quad(lambda s: 4 * np.pi * s * s * dens(s), 0, R)[0]
    g = lambda s: GC * menc(s) / s ** 2  # noqa: E731

    def inner(th, s):
        d = np.sqrt(s * s + r * r - 2 * s * r * np.cos(th))
        return dens(d) * g(s) * np.cos(th) * 2 * np.pi * s * s * np.sin(th)
    val = integrate.dblquad(inner, max(1e-9, r - R), r + R, 0, np.pi, epsabs=1e-11, epsrel=1e-9)[0]
    return val / norm2


@pytest.mark.parametrize("cs", [0, 1])
def test_reference_force_is_two_cloud_force(oracle, cs):
    a = 2.0
    for r in (0.3, 0.8, 1.2, 1.9, 2.5):
        assert abs(oracle.reference_force(cs, r, a) - _two_sphere_force(cs, r, a)) < 2e-5 * GC / a ** 2, r


@pytest.mark.parametrize("cs", [0, 1])
def test_sr_pair_factor_is_two_cloud_residual(oracle, cs):
    # F r = -G/r^2 + R(r): the Newtonian force less the two-cloud force,
    # found by integrating the clouds (not from the closed form); zero from
    # r = a on without softening; the softened force away from the clouds
    a = 2.0
    r = np.array([0.3, 0.8, 1.2, 1.9, 2.0, 2.5, 3.1])
    F = oracle.sr_pair_factor(cs, a, 0.0, r * r)
    for k, x in enumerate(r):
        want = -GC / x ** 2 + (_two_sphere_force(cs, x, a) if x < a else GC / x ** 2)
        assert abs(F[k] * x - want) < 2e-5 * GC / a ** 2, x
    assert np.all(F[r >= a] == 0.0) or np.max(np.abs(F[r >= a] * r[r >= a] ** 3)) < 1e-12 * GC
    eps = 0.4
    Fe = oracle.sr_pair_factor(cs, a, eps, r * r)
    far = r >= a
    assert np.allclose(Fe[far] * r[far], GC / r[far] ** 2 - GC / (r[far] ** 2 + eps ** 2), rtol=1e-12, atol=0)
    # the Newtonian singularity at small r, and no self pair
    tiny = np.array([1e-6, 1e-4]) * a
    assert np.allclose(oracle.sr_pair_factor(cs, a, 0.0, tiny ** 2) * tiny ** 3, -GC, rtol=1e-3)
    assert oracle.sr_pair_factor(cs, a, 0.0, np.zeros(1))[0] == 0.0


def test_sr_table_vanishes_beyond_diameter(oracle):
    # with eps=0 and r >= a the reference force is Newtonian: T = 0
    rc, a = 3.0, 2.0
    t, d2 = oracle.sr_table(rc, a, 0.0, 1)
    r = np.sqrt(np.arange(500) * d2)
    assert np.all(np.abs(t[r >= a * 1.0001]) < 1e-6 * np.abs(t[1:]).max())
    assert t[0] == 0.0


@pytest.mark.parametrize("cs,eps", [(0, 0.4), (1, 0.4), (1, 1.1)])
def test_sr_reference_table_with_softening(oracle, cs, eps):
    # initSRForceTable (source/p3mMethod.cpp:275-294) divides by sqrt(r^2+eps^2)
    # rather than r.  Each entry is therefore the direct pair factor F (pinned
    # above by the two-cloud integration and the softened far field) times
    # r/sqrt(r^2+eps^2): a dropped softening term, a wrong sign of R or an
    # off-by-one in r^2 = i * delta^2 fails this.
    rc, a = 3.0, 2.0
    t, d2 = oracle.sr_table(rc, a, eps, cs)
    assert abs(d2 - np.float32(rc * rc) / np.float32(499)) <= 1e-7 * d2
    r2 = np.arange(500, dtype=np.float64) * np.float64(np.float32(d2))
    r = np.sqrt(r2)
    F = oracle.sr_pair_factor(cs, a, eps, r2[1:])
    want = F * r[1:] / np.sqrt(r2[1:] + eps * eps)
    assert t[0] == 0.0
    assert np.allclose(t[1:], want, rtol=2e-5, atol=2e-6 * np.abs(want).max())


def _brute_sr(pos, m, rc, a, cs):
    d = pos[:, None, :].astype(np.float64) - pos[None, :, :]
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    R = np.vectorize(lambda x: _ref_force64(cs, x, a))(np.where(np.isinf(r), 1.0, r))
    f = np.where(r < rc, (-GC / r ** 2 + R) / r, 0.0)
    return (f[:, :, None] * m[None, :, None] * d).sum(1)


def _ref_force64(cs, r, a):
    if cs == 0:
        if r >= a:
            return GC / r ** 2
        return GC / a ** 2 * (8 * r / a - 9 * r * r / a ** 2 + 2 * (r / a) ** 4)
    u = 2 * r / a
    if u <= 1:
        return GC / (35 * a * a) * (224 * u - 224 * u ** 3 + 70 * u ** 4 + 48 * u ** 5 - 21 * u ** 6)
    if u <= 2:
        return GC / (35 * a * a) * (12 / u ** 2 - 224 + 896 * u - 840 * u ** 2 + 224 * u ** 3 + 70 * u ** 4
                                    - 48 * u ** 5 + 7 * u ** 6)
    return GC / r ** 2


@pytest.mark.parametrize("cs", [0, 1])
@pytest.mark.parametrize("ysort", [True, False])
def test_sr_matches_brute_force(oracle, cs, ysort):
    r = _rng(10)
    pos = r.uniform(0.5, 11.5, (400, 3)).astype(np.float32)
    m = r.uniform(0.5, 1.5, 400).astype(np.float32)
    rc, a, H = 3.0, 2.5, 0.5
    box = np.array([6.0, 6.0, 6.0], np.float32)  # original units; chaining cells int(6/1.5)=4
    brute = _brute_sr(pos, m, rc, a, cs)
    direct, pairs = oracle.p3m_short_range(pos, m, box, rc * H, H, rc, a, 0.0, cs, use_table=False, ysort=ysort)
    assert np.allclose(direct, brute, rtol=2e-4, atol=2e-5 * np.abs(brute).max())
    d = pos[:, None, :] - pos[None, :, :]
    rr = np.linalg.norm(d.astype(np.float64), axis=2)
    assert pairs == int(((rr < rc) & (rr > 0)).sum())


@pytest.mark.parametrize("cs", [0, 1])
def test_sr_table_matches_direct_formula_away_from_zero(oracle, cs):
    # The 500-entry table interpolates linearly in r^2 (p3mMethod.cpp:240-245),
    # which is coarse where T ~ r^-3 near r = 0; on a jittered lattice whose
    # pairs are all further than one cell apart the lerp error is small.
    r = _rng(13)
    g = np.stack(np.meshgrid(*[np.arange(1.0, 11.0, 1.3)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (g + r.uniform(-0.1, 0.1, g.shape)).astype(np.float32)
    m = r.uniform(0.5, 1.5, len(pos)).astype(np.float32)
    rc, a, H = 3.0, 2.5, 0.5
    box = np.array([6.0, 6.0, 6.0], np.float32)
    direct, _ = oracle.p3m_short_range(pos, m, box, rc * H, H, rc, a, 0.0, cs, use_table=False)
    table, _ = oracle.p3m_short_range(pos, m, box, rc * H, H, rc, a, 0.0, cs, use_table=True)
    assert np.linalg.norm(table - direct) / np.linalg.norm(direct) < 3e-3


@pytest.mark.parametrize("cs,eps", [(0, 0.0), (1, 0.0), (1, 0.4)])
def test_sr_accurate_table_matches_brute_force(oracle, cs, eps):
    # the r^3-scaled table follows the direct formula including close pairs
    r = _rng(14)
    pos = r.uniform(0.5, 11.5, (400, 3)).astype(np.float32)
    m = r.uniform(0.5, 1.5, 400).astype(np.float32)
    rc, a, H = 3.0, 2.5, 0.5
    box = np.array([6.0, 6.0, 6.0], np.float32)
    direct, _ = oracle.p3m_short_range(pos, m, box, rc * H, H, rc, a, eps, cs, use_table=0)
    acc, _ = oracle.p3m_short_range(pos, m, box, rc * H, H, rc, a, eps, cs, use_table=2)
    assert np.linalg.norm(acc - direct) / np.linalg.norm(direct) < (3e-5 if eps == 0.0 else 1e-3)
    if eps == 0.0:
        assert np.linalg.norm(direct - _brute_sr(pos, m, rc, a, cs)) / np.linalg.norm(direct) < 1e-4


def test_sr_is_antisymmetric(oracle):
    r = _rng(11)
    pos = r.uniform(0.5, 11.5, (300, 3)).astype(np.float32)
    m = r.uniform(0.5, 1.5, 300).astype(np.float32)
    acc, _ = oracle.p3m_short_range(pos, m, np.array([6.0, 6, 6], np.float32), 1.5, 0.5, 3.0, 2.5, 0.0, 1)
    p = (m[:, None].astype(np.float64) * acc).sum(0)
    assert np.max(np.abs(p)) < 1e-5 * np.abs(m[:, None] * acc).sum()


# --------------------------------------------------------------------------
# direct sum
# --------------------------------------------------------------------------

def test_direct_sum_two_bodies(oracle):
    pos = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    m = np.array([3.0, 5.0], np.float32)
    acc = oracle.direct_sum(pos, m, GC)
    assert np.allclose(acc, [[GC * 5 / 4, 0, 0], [-GC * 3 / 4, 0, 0]])
    acc1 = oracle.direct_sum(pos, m, GC, eps=1.0, kernel=1)
    assert np.allclose(acc1[0, 0], GC * 5 * 2 / 5 ** 1.5)
    acc2 = oracle.direct_sum(pos, m, GC, eps=1.0, kernel=2, targets=[1])
    assert np.allclose(acc2[0, 0], -GC * 3 / 5)


@pytest.mark.parametrize("cs,a,eps", [(0, 2.5, 0.0), (1, 2.5, 0.0), (1, 3.0, 0.0), (1, 2.5, 0.4), (0, 3.5, 0.0)])
def test_sr_log_table_follows_the_pair_factor(oracle, cs, a, eps):
    # the bit-binned table against the closed-form pair factor in FP64: relative
    # error bounded where F is not near a zero, zero from the cutoff (or, with
    # no softening and a <= rc, from r = a) on, and the value of the first bin
    # for closer pairs
    rc = 3.0
    tab, c, lo = oracle.sr_table_log(rc, a, eps, cs)
    assert tab.shape == (4096, 2)
    unscaled = eps == 0.0 and a <= rc
    assert c == (1.0 if unscaled else np.float32(1 / 9))
    hi = lo + 4095 * 65536
    if unscaled:  # the zero bin starts at the bin boundary at or above a^2
        assert hi - 65536 < np.float32(a * a).view(np.uint32) <= hi
    else:  # ... or at the cutoff s = 1
        assert hi == np.float32(1.0).view(np.uint32)

    def exact(r):
        return (-GC / (r * r + eps * eps) + np.vectorize(lambda x: _ref_force64(cs, x, a))(r)) / r

    def lookup(r2):
        s = (np.float32(r2) * np.float32(c)).astype(np.float32)
        b = np.clip(s.view(np.uint32), lo, hi)
        k = (b >> 16) - (lo >> 16)
        y = ((b & 0xFFFF) | 0x3F800000).astype(np.uint32).view(np.float32)
        return tab[k, 0] + tab[k, 1] * y

    r = np.exp(_rng(15).uniform(np.log(1e-2), np.log(0.6 * min(a, rc)), 20000))
    r2 = (r * r).astype(np.float32)
    ref = exact(np.sqrt(r2.astype(np.float64)))
    rr = np.sqrt(r2.astype(np.float64))
    err = np.abs(lookup(r2) - ref)
    # relative to F where the mesh part is small, and to the bare softened
    # Newtonian factor where F itself tends to zero (r -> a)
    assert np.max((err / np.abs(ref))[rr < 0.4 * min(a, rc)]) < 3e-5
    assert np.max(err / (GC / (rr * (rr * rr + eps * eps)))) < 2e-5
    far = np.float32([9.0, 9.5, 40.0, np.inf])
    assert np.all(lookup(far) == 0.0)
    if unscaled:  # F vanishes from r = a on
        assert np.all(lookup(np.float32([a * a * 1.02, 0.5 * (a * a + 9.0)])) == 0.0)
    elif a > rc:  # the force is cut at rc, not before
        assert np.all(lookup(np.float32([8.5, 8.9])) != 0.0)
    first = np.uint32(lo).view(np.float32) / np.float32(c)
    assert np.all(lookup(np.float32([1e-30, 0.0])) == lookup(np.float32([first]))[0])
