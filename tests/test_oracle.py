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


def _m2_grid(N):
    l = _lat(N).astype(np.int64)
    return l[:, None, None] ** 2 + l[None, :, None] ** 2 + l[None, None, :] ** 2


def _frac(x):
    from fractions import Fraction
    return Fraction(x)


@pytest.mark.parametrize("mmin,mmax,bw", [
    ("3.3", "5.45", "1"), ("0.3", "2.2", "1"), ("6.7", "6.9", "1"), ("4.2", "7.6", "1.5"),
    ("2.35", "3.1", "2"), ("5.8", "5.3", "1"), ("9.1", "8.2", "1"), ("0.0", "1.7", "0.5")])
def test_range_mask_integer_bruteforce(mmin, mmax, bw):
    """RANGE (filter.c:130-162): a cell is selected iff mmin - |m| <= bw/2 and
    |m| - mmax <= bw/2, i.e. lo <= |m| <= hi with lo = mmin - bw/2 and
    hi = mmax + bw/2.  With the exact rationals lo, hi and the integer |m|^2
    that is (lo <= 0 or |m|^2 >= lo^2) and |m|^2 <= hi^2, decided without any
    floating point.  The bounds sit off the lattice (no |m|^2 within 1e-6 of
    lo^2 or hi^2), so one-ulp roundings in the oracle cannot move a cell.  A
    swapped difference (m - mmin), a dropped half width or bounds taken as
    strict would each move cells.  Last two cases: mmin > mmax (an empty, then
    a thin, shell) and bw = 0.5."""
    N, L = 24, 60.0
    lo = _frac(mmin) - _frac(bw) / 2
    hi = _frac(mmax) + _frac(bw) / 2
    m2 = _m2_grid(N)
    for v in np.unique(m2):
        assert abs(v - lo * lo) > 1e-6 and abs(v - hi * hi) > 1e-6
    ref = np.zeros(m2.shape, dtype=bool)
    for v in np.unique(m2):
        ok = (lo <= 0 or v >= lo * lo) and v <= hi * hi
        ref |= (m2 == v) & ok
    kmin = float(mmin) * 2 * PI / L
    kmax = float(mmax) * 2 * PI / L
    m, lo_out, hi_out = O.kfilter(N, O.MASK_RANGE, 0.0, float(bw), L, kmin, kmax)
    np.testing.assert_array_equal(m.astype(bool), ref)
    assert (lo_out, hi_out) == (kmin, kmax)  # read, not written


@pytest.mark.parametrize("kkf", [0.0, 2.0, 4.37, 7.5])
def test_range_mask_with_kmin_eq_kmax_is_default(kkf):
    """RANGE with kmin = kmax = k selects the DEFAULT shell of k exactly: the
    two one-sided tests (filter.c:155-158) are |k/kF - m| <= bw/2 (filter.c:57-59)
    split in two, and m - k/kF is exactly -(k/kF - m) in floating point.  This
    is what the n = 2 last leg of DERIVED binning selects (intended meaning of
    filter.c:321-325)."""
    N, L = 24, 60.0
    k = kkf * 2 * PI / L
    for bw in (1.0, 2.0):
        r, _, _ = O.kfilter(N, O.MASK_RANGE, 0.0, bw, L, k, k)
        d, _, _ = O.kfilter(N, O.MASK_DEFAULT, k, bw, L)
        np.testing.assert_array_equal(r, d)


@pytest.mark.parametrize("mmin,mmax,bw", [
    ("3.3", "5.45", "1"), ("0.2", "1.3", "1"), ("6.8", "6.9", "1"), ("4.2", "7.6", "2"),
    ("5.7", "3.2", "1"), ("8.35", "8.1", "1")])
def test_watkinson_kn_mask_interval_overlap(mmin, mmax, bw):
    """WATKINSON_KN (filter.c:219-250): the cube of side bw around m has corner
    norms in [pmin, pmax] with pmin^2 = sum_i (|m_i| - e)^2 and pmax^2 =
    sum_i (|m_i| + e)^2, e = bw/2 (min and max of (m_i -/+ e)^2 per axis; the
    closed form of get_corners is pinned in test_corner_extremes_closed_form).
    For mmin <= mmax the three clauses of filter.c:243-246 say that the
    intervals [pmin, pmax] and [mmin, mmax] overlap: pmin <= mmax and
    pmax >= mmin.  For mmin > mmax they say that [pmin, pmax] contains mmin or
    mmax.  Decided in exact rationals; bounds off every corner norm."""
    from fractions import Fraction
    N, L = 20, 50.0
    e = _frac(bw) / 2
    a, b = _frac(mmin), _frac(mmax)
    l = [abs(int(v)) for v in _lat(N)]
    ref = np.zeros((N, N, N), dtype=bool)
    for i in range(N):
        for j in range(N):
            for k in range(N):
                ax, ay, az = Fraction(l[i]), Fraction(l[j]), Fraction(l[k])
                p2 = (ax - e) ** 2 + (ay - e) ** 2 + (az - e) ** 2
                q2 = (ax + e) ** 2 + (ay + e) ** 2 + (az + e) ** 2
                for t in (a * a, b * b):
                    assert abs(p2 - t) > 1e-6 and abs(q2 - t) > 1e-6
                if a <= b:
                    ref[i, j, k] = p2 <= b * b and q2 >= a * a
                else:
                    ref[i, j, k] = (p2 <= a * a <= q2) or (p2 <= b * b <= q2)
    kmin = float(mmin) * 2 * PI / L
    kmax = float(mmax) * 2 * PI / L
    m, _, _ = O.kfilter(N, O.MASK_WATKINSON_KN, 0.0, float(bw), L, kmin, kmax)
    assert ref.any()
    np.testing.assert_array_equal(m.astype(bool), ref)


def _k3_case(leg0, leg1, k2):
    k = [leg0[1], leg1[1], k2]
    return O.calc_k3_min_and_max(k, [leg0[0], leg1[0], 0.0], [leg0[2], leg1[2], 0.0])


# cos(theta) = 1 (k3 = |k1 - k2|): each candidate of filter.c:300-307 is the
# unique closest pair (a, b), a from leg 1's (kmin, k, kmax), b from leg 2's,
# so the smallest sqrt((a - b)^2) = |a - b| is known by hand; the largest is a
# corner (a function convex in a and b), also by hand.  The last case makes
# the excluded pair (k1, k2) the closest (distance 0): it is not a candidate.
K3_COS1 = [
    # (kmin1, k1, kmax1), (kmin2, k2, kmax2), closest pair, expected lo, hi
    ((1.0, 5.0, 9.0), (2.0, 3.5, 5.1), "k1,kmax2", 0.1, 7.0),
    ((1.0, 2.0, 4.0), (3.0, 4.2, 9.0), "kmax1,k2", 0.2, 8.0),
    ((1.0, 5.0, 9.0), (5.3, 7.0, 12.0), "k1,kmin2", 0.3, 11.0),
    ((4.0, 8.0, 11.0), (1.0, 4.4, 6.0), "kmin1,k2", 0.4, 10.0),
    ((2.0, 6.0, 9.0), (2.5, 12.0, 15.0), "kmin1,kmin2", 0.5, 13.0),
    ((5.0, 9.0, 12.0), (0.5, 2.0, 5.6), "kmin1,kmax2", 0.6, 11.5),
    ((1.0, 2.0, 3.0), (3.7, 6.0, 9.0), "kmax1,kmin2", 0.7, 8.0),
    ((1.0, 2.0, 5.0), (0.1, 3.5, 5.3), "kmax1,kmax2", 0.3, 4.9),
    ((1.0, 5.0, 9.0), (2.0, 5.0, 8.5), "k1,k2 excluded", 0.5, 7.5),
]


@pytest.mark.parametrize("leg0,leg1,pair,lo,hi", K3_COS1, ids=[c[2] for c in K3_COS1])
def test_k3_range_closed_form_cos1(leg0, leg1, pair, lo, hi):
    """calc_k3_min_and_max (filter.c:291-319) at cos(theta) = 1, where every
    candidate k3^2 = a^2 + b^2 - 2ab is (a - b)^2."""
    got_lo, got_hi = _k3_case(leg0, leg1, abs(leg0[1] - leg1[1]))
    assert abs(got_lo - lo) < 1e-12 and abs(got_hi - hi) < 1e-12, (got_lo, got_hi)


@pytest.mark.parametrize("leg0,leg1", [((0.9, 1.0, 1.2), (0.35, 0.5, 0.8)), ((2.0, 3.0, 3.1), (4.0, 4.5, 7.0)),
                                      ((0.01, 0.3, 0.33), (0.2, 0.21, 0.6))])
def test_k3_range_closed_form_cos0_and_minus1(leg0, leg1):
    """cos(theta) = 0: k3^2 = a^2 + b^2, increasing in both, so the range is
    [sqrt(kmin1^2 + kmin2^2), sqrt(kmax1^2 + kmax2^2)]; cos(theta) = -1:
    k3 = a + b, range [kmin1 + kmin2, kmax1 + kmax2]."""
    lo, hi = _k3_case(leg0, leg1, math.sqrt(leg0[1] ** 2 + leg1[1] ** 2))
    assert abs(lo - math.hypot(leg0[0], leg1[0])) < 1e-13 * hi
    assert abs(hi - math.hypot(leg0[2], leg1[2])) < 1e-13 * hi
    lo, hi = _k3_case(leg0, leg1, leg0[1] + leg1[1])
    assert abs(lo - (leg0[0] + leg1[0])) < 1e-12 * hi
    assert abs(hi - (leg0[2] + leg1[2])) < 1e-12 * hi


def test_k3_range_extremes_over_the_rectangle():
    """For any theta, f(a, b) = a^2 + b^2 - 2ab cos(theta) is convex (positive
    semidefinite form), so its maximum over [kmin1, kmax1] x [kmin2, kmax2] is
    at a corner, and the corners are candidates: kmax[2] = sqrt(max of f over
    the rectangle), checked on a dense grid of the rectangle.  For
    cos(theta) <= 0, f increases in a and b: kmin[2] = sqrt(f(kmin1, kmin2)).
    For cos(theta) > 0 the candidates lie in the rectangle: kmin[2] >= its
    minimum."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        k1, k2 = rng.uniform(0.1, 2.0, 2)
        lo1, lo2 = k1 * rng.uniform(0.5, 1.0), k2 * rng.uniform(0.5, 1.0)
        hi1, hi2 = k1 * rng.uniform(1.0, 1.5), k2 * rng.uniform(1.0, 1.5)
        c = rng.uniform(-1.0, 1.0)
        k3 = math.sqrt(k1 * k1 + k2 * k2 - 2 * k1 * k2 * c)
        c_used = (k1 * k1 + k2 * k2 - k3 * k3) / (2 * k1 * k2)
        got_lo, got_hi = _k3_case((lo1, k1, hi1), (lo2, k2, hi2), k3)
        a = np.linspace(lo1, hi1, 61)[:, None]
        b = np.linspace(lo2, hi2, 61)[None, :]
        f = a * a + b * b - 2 * a * b * c_used
        assert abs(got_hi - math.sqrt(f.max())) < 1e-12 * got_hi
        if c_used <= 0:
            assert abs(got_lo - math.sqrt(f[0, 0])) < 1e-12 * got_hi
        else:
            assert got_lo >= math.sqrt(max(f.min(), 0.0)) - 1e-12 * got_hi


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


# ------------------------------------------------------------------ gradient
def _tri_polys(kf, m1, m2, thetas, case, dcos=0.05):
    out = []
    for th in thetas:
        k1, k2 = m1 * kf, m2 * kf
        k3 = O.calc_k3(k1, k2, math.cos(th))
        lo, hi = [k1, k2, 0.0], [k1, k2, 0.0]
        if case in (2, 3):
            lo[2] = O.calc_k3min(k1, k2, th, dcos)
            hi[2] = O.calc_k3max(k1, k2, th, dcos)
        out.append({"n": 3, "k": [k1, k2, k3], "kmin": lo, "kmax": hi})
    return out


def _values(field, polys, bw, case, L, phase=False):
    spec = O.forward_spectrum(field)
    if phase:
        spec = O.calc_phase(spec)
    return np.array([O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], bw, case, L)[0] for p in polys])


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_polygon_gradient_exact_differences(case):
    """The amplitude estimator is a cubic polynomial in the field (B) or a
    quadratic one (P), so Richardson's two-step central difference
    [8 (f(+e) - f(-e)) - (f(+2e) - f(-2e))] / (12 e) equals the derivative
    exactly; only rounding separates it from or_polygon_gradient."""
    N, L = 16, 80.0
    rng = np.random.default_rng(20 + case)
    f = rng.standard_normal((N, N, N))
    kf = 2 * PI / L
    polys = _tri_polys(kf, 3.0, 4.0, [0.6, 2.2], case)
    polys.append({"n": 2, "k": [3 * kf] * 2, "kmin": [3 * kf] * 2, "kmax": [3 * kf] * 2})
    w = rng.standard_normal(len(polys))
    vals, g = O.polygon_gradient(f, polys, 1.0, case, L, w)
    np.testing.assert_allclose(vals, _values(f, polys, 1.0, case, L), rtol=1e-12)
    eps = 0.25
    for x in [(0, 0, 0), (3, 7, 11), (15, 2, 9)]:
        fd = 0.0
        for s, c in ((1, 8.0), (-1, -8.0), (2, -1.0), (-2, 1.0)):
            fp = f.copy()
            fp[x] += s * eps
            fd += c * float(np.dot(w, _values(fp, polys, 1.0, case, L)))
        fd /= 12 * eps
        assert abs(fd - g[x]) <= 1e-9 * np.abs(g).max(), (x, fd, g[x])


def test_polygon_gradient_phase_only_differences():
    """useOnlyPhase: central differences (eps = 1e-4, error O(eps^2)) of the
    phase-only estimator against the pulled-back gradient."""
    N, L = 16, 80.0
    rng = np.random.default_rng(5)
    f = rng.standard_normal((N, N, N))
    kf = 2 * PI / L
    polys = _tri_polys(kf, 3.0, 3.0, [0.9, 2.0], 0) + _tri_polys(kf, 2.5, 4.0, [1.4], 2)
    w = rng.standard_normal(len(polys))
    vals, g = O.polygon_gradient(f, polys, 1.0, 0, L, w, phase=True)
    vals2, _ = O.polygon_gradient(f, polys[-1:], 1.0, 2, L, w[-1:], phase=True)
    np.testing.assert_allclose(vals[:2], _values(f, polys[:2], 1.0, 0, L, phase=True), rtol=1e-12)
    eps = 1e-4
    for x in [(1, 2, 3), (8, 0, 15)]:
        fd = 0.0
        for s in (1, -1):
            fp = f.copy()
            fp[x] += s * eps
            v = np.concatenate([_values(fp, polys[:2], 1.0, 0, L, phase=True),
                                _values(fp, polys[2:], 1.0, 2, L, phase=True)])
            fd += s * float(np.dot(w, v))
        fd /= 2 * eps
        # the GIVEN polygon went through case 0 above: take its own gradient
        _, g0 = O.polygon_gradient(f, polys[:2], 1.0, 0, L, w[:2], phase=True)
        _, g2 = O.polygon_gradient(f, polys[2:], 1.0, 2, L, w[2:], phase=True)
        gx = g0[x] + g2[x]
        assert abs(fd - gx) <= 1e-6 * np.abs(g0 + g2).max(), (x, fd, gx)
    assert np.isfinite(vals2).all()


def test_polygon_gradient_matches_numpy_restatement():
    """Against numpy, differentiated the other way round:
    dS/d delta = sum_i N^3 ifftn(m_i fftn(Q_i)), Q_i = prod_{j != i} D_j."""
    N, L = 24, 60.0
    rng = np.random.default_rng(8)
    f = rng.standard_normal((N, N, N))
    kf = 2 * PI / L
    polys = _tri_polys(kf, 4.0, 5.0, [0.3, 1.7, 2.9], 0)
    w = rng.standard_normal(len(polys))
    _, g = O.polygon_gradient(f, polys, 1.0, 0, L, w)
    dk = np.fft.fftn(f)
    ref = np.zeros_like(f)
    for p, wt in zip(polys, w):
        ms = [O.kfilter(N, O.MASK_DEFAULT, kk, 1.0, L)[0].astype(float) for kk in p["k"]]
        D = [(np.fft.ifftn(m * dk) * N ** 3).real for m in ms]
        C = np.prod([(np.fft.ifftn(m) * N ** 3).real for m in ms], axis=0).sum()
        for i in range(3):
            Q = np.prod([D[j] for j in range(3) if j != i], axis=0)
            ref += wt * L ** 6 * N ** -9.0 / C * (np.fft.ifftn(ms[i] * np.fft.fftn(Q)) * N ** 3).real
    assert np.abs(g - ref).max() <= 1e-11 * np.abs(ref).max()
