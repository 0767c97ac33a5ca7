"""Analytic self-check fields: sums of plane waves, whose bispectrum has a
closed form (README.rst:6 points to such checks, bispectra_tests.pdf).

delta(x) = sum_j A_j cos(2 pi m_j . x / N + phi_j) has exactly the modes
+-m_j, with delta_hat(+-m_j) = N^3 A_j e^{+-i phi_j} / 2 (the unnormalised
forward transform of fft.c:19-37).  The estimator's sum over the grid is
N^3 times the sum over closed lattice triangles of the product of the three
modes (Documentation.tex:124-162), and only triangles of planted modes have a
non-zero product, so
    sum prod D = N^3 * sum over a + b + c = 0 (mod N), a in shell 1 cap S,
                 b in shell 2 cap S, c in shell 3 cap S of d(a) d(b) d(c),
a sum over the few planted modes S.  For one triple m_1 + m_2 + m_3 = 0 on
three distinct shells it is N^3 * 2 (N^3/2)^3 A_1 A_2 A_3 cos(phi_1 + phi_2
+ phi_3).  The polygon count (the normalisation) is the oracle's exact
lattice count.  Checked through the batched path (FP64 to 1e-10 of the
panel's largest |B|, mixed to 1e-5), the reference-shaped FP64 path and the
transform-free check; B(theta) of a k1 = 2 k2 family with planted triples at
chosen angles is non-zero exactly in the theta bins whose k3 shells hold a
planted third mode (the shape of the reference's test panels).  (Phase-only
fields are left out: the rounding-level modes of a plane-wave field would be
normalised to unit amplitude.)
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)


def plane_waves(N, waves):
    """waves: [(m (3 ints), amplitude, phase)] -> float64 N^3 field."""
    i = np.arange(N, dtype=np.float64)
    x, y, z = np.meshgrid(i, i, i, indexing="ij")
    f = np.zeros((N, N, N))
    for m, a, ph in waves:
        f += a * np.cos(2 * PI * (m[0] * x + m[1] * y + m[2] * z) / N + ph)
    return f


def modes(N, waves):
    """{lattice mode (tuple in [-N/2, N/2)): delta_hat}."""
    d = {}
    for m, a, ph in waves:
        for s in (1, -1):
            key = tuple(((s * c + N // 2) % N) - N // 2 for c in m)
            d[key] = d.get(key, 0.0) + N ** 3 * a * np.exp(1j * s * ph) / 2
    return d


def expected(N, L, waves, p, case=0):
    """(B, count) from the planted modes and the oracle's exact count."""
    d = modes(N, waves)
    k, lo, hi = list(p["k"]), list(p["kmin"]), list(p["kmax"])
    masks = []
    for leg in range(3):
        if case == 0:
            masks.append(O.kfilter(N, O.MASK_DEFAULT, k[leg], 1.0, L)[0])
        else:  # GIVEN: the last leg is the k3 range (filter.c:130-162)
            masks.append(O.kfilter(N, O.MASK_DEFAULT, k[leg], 1.0, L)[0] if leg < 2 else
                         O.kfilter(N, O.MASK_RANGE, 0.0, 1.0, L, lo[2], hi[2])[0])
    idx = lambda m: tuple(c % N for c in m)
    S = 0.0
    for a, b in itertools.product(d, d):
        c = tuple(((-(a[t] + b[t]) + N // 2) % N) - N // 2 for t in range(3))
        if c in d and masks[0][idx(a)] and masks[1][idx(b)] and masks[2][idx(c)]:
            S += (d[a] * d[b] * d[c]).real
    count = O.num_polygons(N, 3, k, lo, hi, 1.0, case, L)
    V = L ** 3
    return V ** 2 * N ** -9.0 * N ** 3 * S / (N ** 3 * round(count)), round(count)


def test_single_triple_closed_form(gpu_lib):
    """m1 = (7, 2, 0), m2 = (-3, 5, 1), m3 = -(m1 + m2) on three distinct
    shells: B = V^2 N^-9 * 2 (N^3/2)^3 A1 A2 A3 cos(phi1 + phi2 + phi3) / count."""
    P = gpu_lib
    N, L = 64, 100.0
    kf = 2 * PI / L
    m1, m2 = (7, 2, 0), (-3, 5, 1)
    m3 = tuple(-(a + b) for a, b in zip(m1, m2))
    waves = [(m1, 0.7, 0.3), (m2, 1.1, -1.2), (m3, 0.9, 2.0)]
    f = plane_waves(N, waves)
    k = [math.sqrt(sum(c * c for c in m)) * kf for m in (m1, m2, m3)]
    p = {"n": 3, "k": k, "kmin": [0.0] * 3, "kmax": [0.0] * 3}
    ref, count = expected(N, L, waves, p)
    closed = L ** 6 * N ** -9.0 * 2 * (N ** 3 / 2) ** 3 * 0.7 * 1.1 * 0.9 * math.cos(0.3 - 1.2 + 2.0) / count
    assert abs(ref - closed) <= 1e-13 * abs(closed)
    out = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            b = ctx.batch([p], 1.0, 0)
            v, npg, _ = b.execute()
            b.close()
            out[prec] = v[0]
            assert round(npg[0]) == count
            if prec == P.PREC_FP64:
                plain = ctx.polyspectrum(3, k, [0.0] * 3, [0.0] * 3, 1.0, 0)[0]
                vd, nd, _ = ctx.direct_check([p], 1.0, 0)
                assert abs(plain - closed) <= 1e-10 * abs(closed)
                assert abs(vd[0] - closed) <= 1e-12 * abs(closed) and nd[0] == count
    print("closed form", closed, "fp64", out[P.PREC_FP64], "mixed", out[P.PREC_MIXED])
    assert abs(out[P.PREC_FP64] - closed) <= 1e-10 * abs(closed)
    assert abs(out[P.PREC_MIXED] - closed) <= 1e-5 * abs(closed)


def test_theta_panel_k1_2k2(gpu_lib):
    """k1 = 2 k2 (|m1| = 12, |m2| = 6), 16 theta bins, GIVEN binning with
    dcos = 0.05: triples planted at three angles (cos theta = -m1.m2 / k1 k2 =
    -1, 0, 1, so k3 = |m1 + m2| = 18, sqrt(180), 6) give B(theta) non-zero in
    the bins whose k3 range holds their third mode, each equal to the sum over
    the planted modes (at theta = 0, k3 = k2 and m3 = m2)."""
    P = gpu_lib
    N, L = 64, 100.0
    kf = 2 * PI / L
    m1s = [(12, 0, 0), (0, 12, 0), (0, 0, 12)]
    m2s = [(6, 0, 0), (0, 0, 6), (0, 0, -6)]
    waves = []
    rng = np.random.default_rng(3)
    for m1, m2 in zip(m1s, m2s):
        m3 = tuple(-(a + b) for a, b in zip(m1, m2))
        for m in (m1, m2, m3):
            waves.append((m, float(rng.uniform(0.5, 1.5)), float(rng.uniform(-PI, PI))))
    f = plane_waves(N, waves)
    k1, k2 = 12 * kf, 6 * kf
    polys = []
    for i in range(16):
        th = PI * i / 16
        k3 = math.sqrt(k1 * k1 + k2 * k2 - 2 * k1 * k2 * math.cos(th))
        lo = math.sqrt(k1 * k1 + k2 * k2 - 2 * k1 * k2 * min(1., math.cos(th) + 0.025))
        hi = math.sqrt(k1 * k1 + k2 * k2 - 2 * k1 * k2 * max(-1., math.cos(th) - 0.025))
        polys.append({"n": 3, "k": [k1, k2, k3], "kmin": [k1, k2, lo], "kmax": [k1, k2, hi]})
    ref = np.array([expected(N, L, waves, p, case=2)[0] for p in polys])
    assert np.count_nonzero(ref) >= 2
    res = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            b = ctx.batch(polys, 1.0, 2)
            res[prec], _, _ = b.execute()
            b.close()
    scale = np.abs(ref).max()
    print("B(theta) closed form:", np.round(ref / scale, 4))
    assert np.abs(res[P.PREC_FP64] - ref).max() <= 1e-10 * scale
    assert np.abs(res[P.PREC_MIXED] - ref).max() <= 1e-5 * scale
