"""Pins of the CPU FP64 oracle (oracle/ps_oracle.c) on the CPU.

Each check is chosen so that a plausible mistake in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails one of them:
  * FFT: brute-force DFT definition, numpy.fft, Parseval, impulse, plane wave;
  * masks: integer brute force of each criterion, closed-form corner extremes,
    golden popcounts / index sums / derived kmin-kmax from the independent
    numpy restatement (tests/golden/make_golden.py);
  * estimator: the direct lattice-polygon sum (no FFT at all), integer
    polygon counts, scaling laws, periodic-shift invariance, the exact
    phase-only power spectrum, and the golden values.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "estimator_small.json")))
PI = math.acos(-1.0)


def dft_bruteforce(x, sign):
    n = len(x)
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return W @ x


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 9, 12, 15, 16, 24, 27, 48])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_definition(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = dft_bruteforce(x, sign)
    np.testing.assert_allclose(O.fft1d(x, sign), ref, rtol=0, atol=1e-13 * max(1, np.abs(ref).max()))


@pytest.mark.parametrize("n", [4, 6, 8])
def test_fft3d_definition(n):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    i = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(i, i) / n)
    ref = np.einsum("ai,bj,ck,ijk->abc", W, W, W, x)
    np.testing.assert_allclose(O.fft3d(x, -1), ref, atol=1e-12 * np.abs(ref).max())
    Wb = np.conj(W)
    refb = np.einsum("ai,bj,ck,ijk->abc", Wb, Wb, Wb, x)
    np.testing.assert_allclose(O.fft3d(x, +1), refb, atol=1e-12 * np.abs(refb).max())


@pytest.mark.parametrize("n", [16, 24, 32, 64])
def test_fft3d_numpy_parseval(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n))
    f = O.fft3d(x, -1)
    ref = np.fft.fftn(x)
    assert np.abs(f - ref).max() < 1e-12 * np.abs(ref).max()
    # Parseval: sum |F|^2 = N^3 sum |x|^2 ; backward(forward(x)) = N^3 x
    assert abs((np.abs(f) ** 2).sum() - n ** 3 * (x ** 2).sum()) < 1e-10 * n ** 3 * (x ** 2).sum()
    assert np.abs(O.fft3d(f, +1) - n ** 3 * x).max() < 1e-10 * n ** 3


def test_fft3d_impulse_and_plane_wave():
    n = 12
    x = np.zeros((n, n, n), complex)
    x[2, 5, 7] = 1.0
    f = O.fft3d(x, -1)
    a, b, c = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    np.testing.assert_allclose(f, np.exp(-2j * np.pi * (2 * a + 5 * b + 7 * c) / n), atol=1e-13)
    # a plane wave back-transforms to a single spike of height N^3
    g = O.fft3d(np.exp(-2j * np.pi * (1 * a + 3 * b + 10 * c) / n), +1)
    assert abs(g[1, 3, 10] - n ** 3) < 1e-9
    g[1, 3, 10] = 0
    assert np.abs(g).max() < 1e-9


# ------------------------------------------------------------------ masks
def _lat(n):
    i = np.arange(n)
    return np.where(i < n // 2, i, i - n)


def test_default_mask_integer_bruteforce():
    """|k/kF - |m|| <= bw/2 with k/kF = c + 0.25 (never on a boundary) is
    (c - 0.25)^2 <= m^2 <= (c + 0.75)^2 in integers."""
    N, L = 24, 60.0
    for c in (1, 3, 5, 8):
        kkf = c + 0.25
        k = kkf * 2 * PI / L
        m, _, _ = O.kfilter(N, O.MASK_DEFAULT, k, 1.0, L)
        l = _lat(N)
        m2 = l[:, None, None] ** 2 + l[None, :, None] ** 2 + l[None, None, :] ** 2
        ref = (16 * m2 >= (4 * c - 1) ** 2) & (16 * m2 <= (4 * c + 3) ** 2)
        np.testing.assert_array_equal(m.astype(bool), ref)


def test_corner_extremes_closed_form():
    """min/max of the 8 corner norms = sqrt(sum (|m_i| -/+ e)^2) (|m_i| >= e)."""
    rng = np.random.default_rng(0)
    for _ in range(200):
        mx, my, mz = rng.integers(-20, 21, 3)
        bw = float(rng.choice([1.0, 2.0, 0.5]))
        e = bw / 2
        lo, hi = O.get_corners(int(mx), int(my), int(mz), bw)
        a = [abs(int(v)) for v in (mx, my, mz)]
        if min(a) >= e:
            assert lo == math.sqrt(sum((v - e) ** 2 for v in a))
        assert hi == math.sqrt(sum((v + e) ** 2 for v in a))


def test_watkinson_mask_criterion():
    """WATKINSON: the cube of side bw around m straddles the sphere k/kF."""
    N, L, bw = 16, 40.0, 1.0
    kkf = 4.3
    m, _, _ = O.kfilter(N, O.MASK_WATKINSON, kkf * 2 * PI / L, bw, L)
    l = _lat(N)
    e = bw / 2
    a = np.abs(l).astype(float)
    lo = np.sqrt((a - e)[:, None, None] ** 2 + (a - e)[None, :, None] ** 2 + (a - e)[None, None, :] ** 2)
    hi = np.sqrt((a + e)[:, None, None] ** 2 + (a + e)[None, :, None] ** 2 + (a + e)[None, None, :] ** 2)
    ref = (lo <= kkf) & (kkf <= hi)
    np.testing.assert_array_equal(m.astype(bool), ref)


@pytest.mark.parametrize("ent", GOLD["masks"], ids=lambda e: f"kind{e['kind']}-k{e['k']:.3f}")
def test_mask_golden(ent):
    m, lo, hi = O.kfilter(ent["N"], ent["kind"], ent["k"], ent["bw"], ent["L"], ent["kmin"], ent["kmax"])
    assert int(m.sum()) == ent["count"]
    assert int(np.flatnonzero(m.ravel()).sum()) == ent["index_sum"]
    if "kmin_out" in ent:
        assert lo == float.fromhex(ent["kmin_out"])
        assert hi == float.fromhex(ent["kmax_out"])


def test_k3_range_degenerate_and_monotone():
    k = [0.3, 0.5, 0.0]
    k[2] = O.calc_k3(k[0], k[1], math.cos(1.2))
    lo, hi = O.calc_k3_min_and_max(k, k, k)
    assert abs(lo - k[2]) < 1e-14 and abs(hi - k[2]) < 1e-14
    lo, hi = O.calc_k3_min_and_max(k, [0.29, 0.48, 0], [0.31, 0.52, 0])
    assert lo < k[2] < hi
    # kvectors.c:233-247 clamps cos at +-1
    assert O.calc_k3min(0.4, 0.4, 0.0, 0.05) == 0.0
    assert abs(O.calc_k3max(0.4, 0.4, math.pi, 0.05) - 0.8) < 1e-15


# ------------------------------------------------------------------ estimator
def _field():
    N = GOLD["N"]
    return np.array(GOLD["field"], dtype=np.float32).reshape(N, N, N)


@pytest.mark.parametrize("c", GOLD["cases"], ids=lambda c: f"n{c['n']}-case{c['case']}-{c['k'][0]:.3f}")
def test_estimator_golden(c):
    N, L = GOLD["N"], GOLD["L"]
    spec = O.forward_spectrum(_field())
    v, lo, hi, sums = O.polyspectrum(spec, c["n"], c["k"], c["kmin"], c["kmax"], c["bw"], c["case"], L)
    assert abs(v - c["value"]) <= 1e-11 * abs(c["value"])
    assert abs(sums[1].real / N ** 3 - c["npolygons"]) < 1e-8
    n = c["n"]
    assert [float(x).hex() for x in lo[:n]] == c["kmin_out"][:n]
    assert [float(x).hex() for x in hi[:n]] == c["kmax_out"][:n]
    npg = O.num_polygons(N, n, c["k"], lo, hi, c["bw"], c["case"], L)
    assert abs(npg - c["npolygons"]) < 1e-8


def test_estimator_equals_direct_polygon_sum():
    """B = V^2 N^-9 S/C with S the direct sum over closed lattice triangles."""
    N, L = 16, 100.0
    rng = np.random.default_rng(9)
    f = rng.standard_normal((N, N, N))
    spec = O.forward_spectrum(f)
    kf = 2 * PI / L
    for (m1, m2, th) in ((3.0, 3.0, 2.0), (2.5, 4.0, 1.1), (4.0, 5.0, 2.7)):
        k = [m1 * kf, m2 * kf, O.calc_k3(m1 * kf, m2 * kf, math.cos(th))]
        v, _, _, sums = O.polyspectrum(spec, 3, k, k, k, 1.0, 0, L)
        masks = [O.kfilter(N, O.MASK_DEFAULT, kk, 1.0, L)[0] for kk in k]
        S, C = O.direct_polygon_sum(spec, masks)
        assert C > 0
        ref = L ** 6 * N ** -9.0 * S.real / C
        assert abs(v - ref) < 1e-12 * abs(ref)
        assert abs(sums[1].real / N ** 3 - C) < 1e-8 * C   # exact integer count
        assert abs(sums[0].imag) < 1e-9 * abs(sums[0].real)
    # n = 2
    k = [3.0 * kf, 3.0 * kf, 0]
    v, _, _, sums = O.polyspectrum(spec, 2, k, k, k, 1.0, 0, L)
    m = O.kfilter(N, O.MASK_DEFAULT, k[0], 1.0, L)[0]
    S, C = O.direct_polygon_sum(spec, [m, m])
    assert C == int(m.sum())
    assert abs(v - L ** 3 * N ** -6.0 * S.real / C) < 1e-12 * abs(v)


def test_estimator_scaling_and_shift():
    N, L = 16, 80.0
    rng = np.random.default_rng(4)
    f = rng.standard_normal((N, N, N))
    kf = 2 * PI / L
    k = [3 * kf, 4 * kf, O.calc_k3(3 * kf, 4 * kf, math.cos(1.9))]
    b1 = O.polyspectrum(O.forward_spectrum(f), 3, k, k, k, 1.0, 0, L)[0]
    b3 = O.polyspectrum(O.forward_spectrum(2.5 * f), 3, k, k, k, 1.0, 0, L)[0]
    assert abs(b3 - 2.5 ** 3 * b1) < 1e-11 * abs(b3)
    bs = O.polyspectrum(O.forward_spectrum(np.roll(f, (3, -5, 7), (0, 1, 2))), 3, k, k, k, 1.0, 0, L)[0]
    assert abs(bs - b1) < 1e-11 * abs(b1)
    p1 = O.polyspectrum(O.forward_spectrum(f), 2, [k[0]] * 2, [k[0]] * 2, [k[0]] * 2, 1.0, 0, L)[0]
    p2 = O.polyspectrum(O.forward_spectrum(-3 * f), 2, [k[0]] * 2, [k[0]] * 2, [k[0]] * 2, 1.0, 0, L)[0]
    assert abs(p2 - 9 * p1) < 1e-11 * abs(p2)


def test_phase_only_power_spectrum_exact():
    """|p| = 1 and p(-m) = conj p(m) give P = V N^-6 exactly."""
    N, L = 16, 70.0
    rng = np.random.default_rng(5)
    spec = O.calc_phase(O.forward_spectrum(rng.standard_normal((N, N, N))))
    kf = 2 * PI / L
    for m in (2.0, 4.5, 6.0):
        k = [m * kf] * 2
        v = O.polyspectrum(spec, 2, k, k, k, 1.0, 0, L)[0]
        assert abs(v - L ** 3 * N ** -6.0) < 1e-12 * L ** 3 * N ** -6.0


def test_num_polygons_counts_integer():
    N, L = 16, 100.0
    kf = 2 * PI / L
    for case in (0, 1, 2, 3):
        k = [3 * kf, 3 * kf, O.calc_k3(3 * kf, 3 * kf, math.cos(2.0))]
        lo = [k[0], k[1], O.calc_k3min(k[0], k[1], 2.0, 0.05)]
        hi = [k[0], k[1], O.calc_k3max(k[0], k[1], 2.0, 0.05)]
        c = O.num_polygons(N, 3, k, lo, hi, 1.0, case, L)
        assert abs(c - round(c)) < 1e-8 and c > 0
