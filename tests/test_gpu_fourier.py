"""The Fourier-space evaluation of a batch (METHOD_FOURIER, opt-in) against
the direct method and the oracle.

Per group of polygons sharing their first legs, the product of those legs is
forward transformed once and every last leg becomes a sum over its shell
(Parseval); power spectra are shell sums of |H|^2.  The identity is exact, so
FP64 agrees with the oracle to 1e-10 relative, with the direct method to
1e-11; the mixed-precision bar is 1e-5 of the sweep's largest |value|.
Covered: triangles with squared and distinct prefix legs, both power-spectrum
kinds (identical legs, different legs), every binning case including the
corner-tested WATKINSON shells, shells reaching the Nyquist planes, several
slots, and the split of a batch into arena waves."""
import math
import os

import numpy as np
import pytest

import oracle as O
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)


def _tris(kf, pairs, thetas, case, dcos=0.05, slot=0):
    out = []
    for (m1, m2) in pairs:
        k1, k2 = m1 * kf, m2 * kf
        for th in thetas:
            k3 = O.calc_k3(k1, k2, math.cos(th))
            lo, hi = [k1, k2, 0.0], [k1, k2, 0.0]
            if case in (2, 3):
                lo[2] = O.calc_k3min(k1, k2, th, dcos)
                hi[2] = O.calc_k3max(k1, k2, th, dcos)
            out.append({"n": 3, "slot": slot, "k": [k1, k2, k3], "kmin": lo, "kmax": hi})
    return out


def _pk(kf, ms, slot=0, other=None):
    out = []
    for m in ms:
        k = m * kf
        k2 = (other if other is not None else m) * kf
        out.append({"n": 2, "slot": slot, "k": [k, k2], "kmin": [k, k2], "kmax": [k, k2]})
    return out


def _run(P, N, L, prec, fields, polys, case, method, params=None):
    with P.Context(N, L, prec, nslots=len(fields)) as ctx:
        for s, f in enumerate(fields):
            ctx.load_field(f, slot=s)
        p = {"grid_cap": 1024, "chunk_tiles": 4, "g2_div": 1, "method": method}
        if params:
            p.update(params)
        b = ctx.batch(polys, 1.0, case, params=p)
        v, npg, sums = b.execute()
        v2, _, _ = b.execute()
        np.testing.assert_array_equal(v, v2)  # graph replay
        b.close()
    return v, npg


def _oracle(fields, polys, case, L, every=1):
    out = {}
    specs = [O.forward_spectrum(f) for f in fields]
    for i, p in enumerate(polys):
        if i % every:
            continue
        out[i] = O.polyspectrum(specs[p.get("slot", 0)], p["n"], p["k"], p["kmin"], p["kmax"], 1.0, case, L)[0]
    return out


@pytest.mark.parametrize("N,L,case", [(32, 100.0, 0), (48, 120.0, 2), (64, 100.0, 1), (32, 90.0, 3)])
def test_fourier_fp64_vs_oracle_and_direct(gpu_lib, N, L, case):
    P = gpu_lib
    f = lognormal_field(N, 4000 + N + case)
    kf = 2 * PI / L
    th = [PI * i / 7 for i in range(7)]
    polys = _tris(kf, [(4.0, 4.0), (3.0, 5.0), (6.5, 2.5)], th, case)
    polys += _pk(kf, [1, 2, 5, 9])                  # identical legs (or not, by binning)
    polys += _pk(kf, [3, 4], other=3.4)             # different legs
    vf, nf = _run(P, N, L, P.PREC_FP64, [f], polys, case, P.METHOD_FOURIER)
    vd, nd = _run(P, N, L, P.PREC_FP64, [f], polys, case, P.METHOD_DIRECT)
    tri = np.array([p["n"] == 3 for p in polys])
    ok = np.isfinite(vd)
    # sums that cancel are compared relative to the sweep's largest value
    np.testing.assert_allclose(vf[tri & ok], vd[tri & ok], rtol=1e-10, atol=1e-12 * np.abs(vd[tri & ok]).max())
    np.testing.assert_allclose(vf[~tri], vd[~tri], rtol=1e-11)
    np.testing.assert_allclose(nf, nd, rtol=1e-12, atol=1e-9)
    for i, ref in _oracle([f], polys, case, L, every=3).items():
        if np.isfinite(ref) and ref != 0:
            assert abs(vf[i] - ref) <= 1e-10 * abs(ref), (i, vf[i], ref)


def test_fourier_mixed_vs_fp64(gpu_lib):
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 99, sigma=0.8, alpha=1.5)
    kf = 2 * PI / L
    th = [PI * i / 12 for i in range(12)]
    polys = _tris(kf, [(7.96, 7.96), (5.0, 9.0), (2.0, 2.0)], th, 2) + _pk(kf, range(1, 20))
    v64, n64 = _run(P, N, L, P.PREC_FP64, [f], polys, 2, P.METHOD_FOURIER)
    v32, n32 = _run(P, N, L, P.PREC_MIXED, [f], polys, 2, P.METHOD_FOURIER)
    np.testing.assert_array_equal(np.round(n32), np.round(n64))
    tri = slice(0, 36)
    err_t = np.abs(v32[tri] - v64[tri]).max() / np.abs(v64[tri]).max()
    err_p = (np.abs(v32[36:] - v64[36:]) / np.abs(v64[36:])).max()
    print("mixed Fourier: bispectrum", err_t, "power", err_p)
    assert err_t < 1e-5 and err_p < 1e-5


def test_fourier_nyquist_shells(gpu_lib):
    """Shells that reach |m| = N/2: the kz = N/2 plane enters with weight 1."""
    P = gpu_lib
    N, L = 32, 2 * PI  # k_F = 1
    f = lognormal_field(N, 7)
    polys = _tris(1.0, [(15.5, 15.5), (14.0, 12.0)], [PI / 3, 2.5], 0) + _pk(1.0, [15, 16, 20])
    vf, nf = _run(P, N, L, P.PREC_FP64, [f], polys, 0, P.METHOD_FOURIER)
    vd, nd = _run(P, N, L, P.PREC_FP64, [f], polys, 0, P.METHOD_DIRECT)
    np.testing.assert_allclose(nf, nd, rtol=1e-12, atol=1e-9)
    # a shell pair with no closing triangle has no polygons: 0/0 in both methods
    np.testing.assert_array_equal(np.isnan(vf), np.isnan(vd))
    top = np.nanmax(np.abs(vd))
    ok = np.abs(vd) > 1e-8 * top
    np.testing.assert_allclose(vf[ok], vd[ok], rtol=1e-10)
    for i, ref in _oracle([f], polys, 0, L).items():
        if ref != 0 and np.isfinite(ref):
            assert abs(vf[i] - ref) <= 1e-10 * abs(ref) + 1e-12 * top, (i, vf[i], ref)


def test_fourier_slots_and_waves(gpu_lib):
    """Two field slots, and the batch split into several arena waves (each
    wave reuses the buffers): identical values to the single-wave plan."""
    P = gpu_lib
    N, L = 48, 100.0
    fields = [lognormal_field(N, 11), lognormal_field(N, 12)]
    kf = 2 * PI / L
    th = [0.3, 1.2, 2.6]
    polys = []
    for s in (0, 1):
        polys += _tris(kf, [(4.0, 4.0), (3.0, 6.0), (7.0, 7.0)], th, 2, slot=s) + _pk(kf, [2, 3, 4], slot=s)
    v1, n1 = _run(P, N, L, P.PREC_FP64, fields, polys, 2, P.METHOD_FOURIER)
    os.environ["PS_ARENA_MB"] = "1"
    try:
        v2, n2 = _run(P, N, L, P.PREC_FP64, fields, polys, 2, P.METHOD_FOURIER)
    finally:
        del os.environ["PS_ARENA_MB"]
    np.testing.assert_array_equal(v2, v1)
    np.testing.assert_array_equal(n2, n1)
    for i, ref in _oracle(fields, polys, 2, L, every=4).items():
        assert abs(v1[i] - ref) <= 1e-10 * abs(ref), (i, v1[i], ref)
