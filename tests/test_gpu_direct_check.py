"""The transform-free check (ps_direct_check) and, through it, the batched
estimator at the benchmark's full 512^3 size.

ps_direct_check sums prod H(m_i) over the closed lattice polygons
m_1 + ... + m_n = 0 (mod N) of the shells, on the device, with no transform.
With the unnormalised transforms, sum_x prod D_i = N^3 * that sum and
sum_x prod I_i = N^3 * the number of polygons, so its values equal the
estimator's (Documentation.tex:124-162) by a different route.

* Small grids: against the oracle's direct lattice sum with the oracle's own
  masks (counts exact, sums to 1e-12) and against the oracle estimator
  (values to 1e-10, counts exact) in every binning, n = 2 and 3, phase-only,
  and more third-leg shells than one pass of the kernel tests.
* 512^3, the benchmark sweep's field: the batched FP64 estimator equals the
  direct check to 1e-10 relative (measured 1.3e-15), its counts round to the
  direct counts, and the mixed one is within 1e-5 of each k1 = k2 group's
  largest |B|.
"""
import math

import numpy as np
import pytest

import oracle as O
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)


def _tri(kf, m1, m2, theta, case, dcos=0.05):
    k1, k2 = m1 * kf, m2 * kf
    k = [k1, k2, O.calc_k3(k1, k2, math.cos(theta))]
    kmin, kmax = [k1, k2, 0.0], [k1, k2, 0.0]
    if case in (2, 3):  # run.c:68-81
        kmin[2] = O.calc_k3min(k1, k2, theta, dcos)
        kmax[2] = O.calc_k3max(k1, k2, theta, dcos)
    return {"n": 3, "k": k, "kmin": kmin, "kmax": kmax}


def _pk(kf, m):
    k = m * kf
    return {"n": 2, "k": [k, k, 0.0], "kmin": [k, k, 0.0], "kmax": [k, k, 0.0]}


def test_direct_check_equals_oracle_lattice_sum(gpu_lib):
    """Counts and sums against the oracle's direct sum over its own masks."""
    P = gpu_lib
    N, L = 24, 100.0
    f = lognormal_field(N, 31, sigma=0.8, alpha=1.5)
    spec = O.forward_spectrum(f)
    kf = 2 * PI / L
    polys = [_tri(kf, 3.0, 3.0, th, 0) for th in (0.0, 0.9, 2.0, 3.0)]
    polys += [_tri(kf, 2.5, 4.0, 1.1, 0), _tri(kf, 7.0, 6.0, 2.6, 0), _pk(kf, 3.0), _pk(kf, 9.0)]
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        v, npg, sums = ctx.direct_check(polys, 1.0, 0)
    for i, p in enumerate(polys):
        n = p["n"]
        masks = [O.kfilter(N, O.MASK_DEFAULT, p["k"][j], 1.0, L)[0] for j in range(n)]
        S, C = O.direct_polygon_sum(spec, masks)
        print(i, n, C, npg[i], S.real, sums[i, 0] / N ** 3)
        assert C > 0 and npg[i] == C
        assert sums[i, 1] == float(N ** 3) * C
        assert abs(sums[i, 0] - N ** 3 * S.real) <= 1e-12 * N ** 3 * abs(S)
        ref = L ** (3 * (n - 1)) * N ** (-3.0 * n) * S.real / C
        assert abs(v[i] - ref) <= 1e-12 * abs(ref)


@pytest.mark.parametrize("case", [0, 1, 2, 3])
@pytest.mark.parametrize("phase", [0, 1])
def test_direct_check_equals_oracle_estimator(gpu_lib, case, phase):
    """Every binning, bispectra (20 theta bins of one k1, k2 pair: two passes
    of third-leg shells) and power spectra, against the oracle estimator."""
    P = gpu_lib
    N, L = 32, 100.0
    f = lognormal_field(N, 77 + case, sigma=0.8, alpha=1.5)
    spec = O.forward_spectrum(f)
    if phase:
        spec = O.calc_phase(spec)
    kf = 2 * PI / L
    polys = [_tri(kf, 5.0, 5.0, PI * i / 20, case) for i in range(20)]
    polys += [_tri(kf, 4.0, 6.5, 1.7, case), _pk(kf, 4.0), _pk(kf, 11.0)]
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        ctx.set_phase_only(phase)
        v, npg, _ = ctx.direct_check(polys, 1.0, case)
    checked = 0
    for i, p in enumerate(polys):
        ref, _, _, rs = O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], 1.0, case, L)
        count = round(rs[1].real / N ** 3)
        assert npg[i] == count, (i, npg[i], rs[1].real / N ** 3)
        if count == 0:
            continue
        checked += 1
        assert abs(v[i] - ref) <= 1e-10 * abs(ref), (i, v[i], ref)
    assert checked >= 18


def test_direct_check_mixed_context(gpu_lib):
    """The check reads the resident spectrum: in a mixed-precision context
    (FP32 spectrum) it agrees with the FP64 one to FP32 rounding, counts
    exact."""
    P = gpu_lib
    N, L = 48, 150.0
    f = lognormal_field(N, 5, sigma=0.8, alpha=1.5)
    kf = 2 * PI / L
    polys = [_tri(kf, 6.0, 6.0, PI * i / 8, 2) for i in range(8)] + [_pk(kf, 6.0)]
    out = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            out[prec] = ctx.direct_check(polys, 1.0, 2)
    v64, n64, _ = out[P.PREC_FP64]
    v32, n32, _ = out[P.PREC_MIXED]
    np.testing.assert_array_equal(n32, n64)
    assert np.max(np.abs(v32 - v64)) <= 1e-5 * np.max(np.abs(v64))


@pytest.mark.parametrize("case", [1, 3])
def test_direct_check_slots_equal_batch(gpu_lib, case):
    """Several field slots in one call (each polygon reads its own slot's
    spectrum), DERIVED and WATKINSON binning, power spectra and bispectra:
    the batched FP64 estimator equals the transform-free sums to 1e-10."""
    P = gpu_lib
    N, L = 64, 100.0
    kf = 2 * PI / L
    polys = []
    for slot in (0, 1):
        polys += [dict(_tri(kf, 6.0, 6.0, PI * i / 7, case), slot=slot) for i in range(7)]
        polys += [dict(_tri(kf, 4.0, 7.0, 1.9, case), slot=slot), dict(_pk(kf, 5.0), slot=slot)]
    with P.Context(N, L, P.PREC_FP64, nslots=2) as ctx:
        ctx.load_field(lognormal_field(N, 61, sigma=0.8, alpha=1.5), slot=0)
        ctx.load_field(lognormal_field(N, 62, sigma=0.8, alpha=1.5), slot=1)
        vd, nd, _ = ctx.direct_check(polys, 1.0, case)
        b = ctx.batch(polys, 1.0, case)
        vb, nb, _ = b.execute()
        b.close()
    closed = nd > 0
    assert closed.sum() >= len(polys) - 2
    np.testing.assert_array_equal(np.round(nb), nd)
    np.testing.assert_allclose(vb[closed], vd[closed], rtol=1e-10, atol=0)
    # the two slots hold different fields
    h = len(polys) // 2
    assert np.all(vd[:h][closed[:h]] != vd[h:][closed[h:]])


def test_config_512_batch_equals_direct_check(gpu_lib):
    """The benchmark sweep's 512^3 field (log-normal, sigma 1, L = 300), k1 = k2
    in {0.1, 0.4, 0.7, 1.0} h/Mpc with 6 theta bins each, GIVEN binning: the
    batched estimator (FFT) against the transform-free sums."""
    P = gpu_lib
    N, L = 512, 300.0
    field = lognormal_field(N, 2000, sigma=1.0, alpha=1.5)
    polys = []
    for k in (0.1, 0.4, 0.7, 1.0):
        for i in range(6):
            th = PI * (5 * i + 2) / 32
            c = math.cos(th)
            k3 = math.sqrt(2 * k * k - 2 * k * k * c)
            lo = math.sqrt(2 * k * k - 2 * k * k * min(1., c + 0.025))
            hi = math.sqrt(2 * k * k - 2 * k * k * max(-1., c - 0.025))
            polys.append({"n": 3, "k": [k, k, k3], "kmin": [k, k, lo], "kmax": [k, k, hi]})
    case = P.BINNING["GIVEN_BINNING"]
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(field)
        vd, nd, _ = ctx.direct_check(polys, 1.0, case)
        b = ctx.batch(polys, 1.0, case)
        v64, n64, _ = b.execute()
        b.close()
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        v32, n32, _ = b.execute()
        b.close()
    rel = np.abs(v64 - vd) / np.abs(vd)
    print("FP64 batched vs direct: max rel", rel.max(), "counts", nd.min(), nd.max())
    assert np.all(nd > 0)
    # the estimator's counts are sums of FP64 transforms: the integers to
    # rounding (measured 5e-16 relative); the direct counts are exact
    np.testing.assert_array_equal(np.round(n64), nd)
    np.testing.assert_array_equal(n32, n64)
    assert np.max(np.abs(n64 - nd) / nd) <= 1e-13
    assert rel.max() <= 1e-10
    for g in range(4):
        s = slice(6 * g, 6 * g + 6)
        scale = np.max(np.abs(vd[s]))
        err = np.max(np.abs(v32[s] - vd[s])) / scale
        print("group", g, "mixed vs direct / group scale", err)
        assert err <= 1e-5
