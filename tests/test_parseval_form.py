"""The arithmetic of the copy-pipelined product pass (k_zc_dma) in its
Parseval form, checked on the CPU against the real-space sum it replaces.

A wavefront holds, per line, two y rows a, b of the prefix product P along z
(Q = P(a) + i P(b)) and the staged half spectra D(a)[k], D(b)[k], k <= Kz, of
a suffix leg (the raw pair (x, y, z, w) = (Re D(a), Im D(a), Re D(b),
Im D(b)) of plane k).  The pass needs S = sum_z P(a) D(a) + P(b) D(b), D the
unnormalised C2R transform of the half spectrum (imaginary parts dropped at
k = 0 and N/2, as a C2R transform does).  With F the forward DFT of Q:

  S = sum_{k <= Kz} x (A.x + B.x) + y (A.y - B.y) + z (A.y + B.y) + w (B.x - A.x)

with A = F(k), B = F(N - k) (B = 0 at k = 0 and N/2, where y = w = 0).  With
magnitude sums the pass also takes the Cauchy-Schwarz bound
sqrt(sum P^2 * N * sum_k w_k (x^2 + y^2 + z^2 + w^2)) >= sum_z |P D|
(w_k = 1 at k = 0, N/2, else 2).  A sign or index slip in either fails here.
"""
import numpy as np
import pytest


def _case(rng, N, Kz):
    pa, pb = rng.standard_normal(N), rng.standard_normal(N)
    da = np.zeros(N // 2 + 1, complex)
    db = np.zeros(N // 2 + 1, complex)
    da[: Kz + 1] = rng.standard_normal(Kz + 1) + 1j * rng.standard_normal(Kz + 1)
    db[: Kz + 1] = rng.standard_normal(Kz + 1) + 1j * rng.standard_normal(Kz + 1)
    return pa, pb, da, db


def _direct(pa, pb, da, db, N):
    # unnormalised C2R (numpy's irfft divides by N; it ignores Im at 0 and N/2)
    Da = np.fft.irfft(da, N) * N
    Db = np.fft.irfft(db, N) * N
    return float(np.sum(pa * Da + pb * Db)), float(np.sum(np.abs(pa * Da) + np.abs(pb * Db))), Da, Db


def _parseval(pa, pb, da, db, N, Kz):
    F = np.fft.fft(pa + 1j * pb)  # forward, e^{-i}, unnormalised
    s = 0.0
    dd = 0.0
    for k in range(Kz + 1):
        edge = k == 0 or 2 * k == N
        x, y, z, w = da[k].real, da[k].imag, db[k].real, db[k].imag
        if edge:
            y = w = 0.0
        A = F[k]
        B = 0.0 if edge else F[N - k]
        s += x * (A.real + np.real(B)) + y * (A.imag - np.imag(B)) + z * (A.imag + np.imag(B)) \
            + w * (np.real(B) - A.real)
        dd += (1.0 if edge else 2.0) * (x * x + y * y + z * z + w * w)
    bound = np.sqrt(np.sum(pa * pa + pb * pb) * N * dd)
    return s, bound


@pytest.mark.parametrize("N,Kz", [(16, 5), (24, 12), (96, 40), (128, 64), (144, 72), (192, 95), (192, 96)])
def test_parseval_sum_equals_real_space_sum(N, Kz):
    rng = np.random.default_rng(N * 1000 + Kz)
    for _ in range(5):
        pa, pb, da, db = _case(rng, N, Kz)
        sd, sabs, _, _ = _direct(pa, pb, da, db, N)
        sp, bound = _parseval(pa, pb, da, db, N, Kz)
        assert abs(sp - sd) <= 1e-10 * sabs, (sp, sd)
        assert bound >= sabs * (1 - 1e-12)


def test_cauchy_schwarz_bound_is_tight_for_proportional_lines():
    """The bound equals sum |P D| when |P| is proportional to |D|, and is
    larger otherwise: it only ever raises the error estimate."""
    rng = np.random.default_rng(7)
    N, Kz = 64, 20
    _, _, da, db = _case(rng, N, Kz)
    _, _, Da, Db = _direct(np.zeros(N), np.zeros(N), da, db, N)
    pa, pb = 3.0 * Da, 3.0 * Db
    sd, sabs, _, _ = _direct(pa, pb, da, db, N)
    sp, bound = _parseval(pa, pb, da, db, N, Kz)
    assert abs(sp - sd) <= 1e-10 * sabs
    assert abs(bound - sabs) <= 1e-9 * sabs
