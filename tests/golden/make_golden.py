"""Writes tests/golden/estimator_small.json: golden vectors for the estimator.

An independent numpy restatement of the reference algorithm on small grids:
  * masks vectorised over the lattice with the same FP64 expressions as
    filter.c:34-289 (evaluated by numpy, element by element, no FMA);
  * C2C transforms from numpy.fft (pocketfft, not the oracle's FFT):
    FFTW_FORWARD = fftn, FFTW_BACKWARD = N^3 * ifftn (fft.c:19-57);
  * complex products and sums as polyspectrum.c:51-252.
Fields are seeded log-normal grids stored in the fixture as float32 values.
Run:  python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np

PI = math.acos(-1.0)
HERE = os.path.dirname(os.path.abspath(__file__))


def lattice(n):
    i = np.arange(n)
    return np.where(i < n // 2, i, i - n)


def grid_m(n):
    m = lattice(n)
    mx, my, mz = np.meshgrid(m, m, m, indexing="ij")
    return mx, my, mz


def corners(mx, my, mz, bw):
    x, y, z = mx.astype(np.float64), my.astype(np.float64), mz.astype(np.float64)
    e = 0.5 * bw
    c = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                c.append((x + sx * e) * (x + sx * e) + (y + sy * e) * (y + sy * e) + (z + sz * e) * (z + sz * e))
    c = np.stack(c)
    return np.sqrt(c.min(axis=0)), np.sqrt(c.max(axis=0))


def mask(n, kind, k, kmin, kmax, bw, L):
    mx, my, mz = grid_m(n)
    m = np.sqrt((mx * mx + my * my + mz * mz).astype(np.float64))
    kkf = k / (2. * PI) * L
    if kind in (0, 1):
        sel = np.abs(kkf - m) <= 0.5 * bw
        if kind == 1:
            lo, hi = corners(mx[sel], my[sel], mz[sel], bw)
            return sel, (lo.min() * 2. * PI / L, hi.max() * 2. * PI / L)
        return sel, None
    mmin = kmin / (2. * PI) * L
    mmax = kmax / (2. * PI) * L
    if kind == 2:
        return ((mmin - m) <= 0.5 * bw) & ((m - mmax) <= 0.5 * bw), None
    pmin, pmax = corners(mx, my, mz, bw)
    if kind == 3:
        return (pmin <= kkf) & (kkf <= pmax), None
    return ((pmax >= mmax) & (pmin <= mmax)) | ((pmax <= mmax) & (pmin >= mmin)) | ((pmax >= mmin) & (pmin <= mmin)), None


def k3_range(k, kmin, kmax):
    c = (k[0] * k[0] + k[1] * k[1] - k[2] * k[2]) / (2. * k[0] * k[1])
    a = [(kmin[0], kmin[1]), (kmin[0], kmax[1]), (kmax[0], kmin[1]), (kmax[0], kmax[1]),
         (k[0], kmax[1]), (kmax[0], k[1]), (k[0], kmin[1]), (kmin[0], k[1])]
    t = [p * p + q * q - 2. * p * q * c for p, q in a]
    return math.sqrt(min(t)), math.sqrt(max(t))


def estimator(spec, n, k, kmin, kmax, bw, case, L):
    N = spec.shape[0]
    kmin = list(kmin)
    kmax = list(kmax)
    accD = np.ones_like(spec)
    accI = np.ones_like(spec)
    for i in range(n):
        if case == 1:
            if i < n - 1:
                sel, (lo, hi) = mask(N, 1, k[i], 0, 0, bw, L)
                kmin[i], kmax[i] = lo, hi
            else:
                if n == 3:
                    kmin[2], kmax[2] = k3_range(k, kmin, kmax)
                else:
                    kmin[n - 1] = kmax[n - 1] = k[n - 1]
                sel, _ = mask(N, 2, 0, kmin[i], kmax[i], bw, L)
        elif case == 2:
            sel, _ = mask(N, 0, k[i], 0, 0, bw, L) if i < n - 1 else mask(N, 2, 0, kmin[i], kmax[i], bw, L)
        elif case == 3:
            sel, _ = mask(N, 3, k[i], 0, 0, bw, L) if i < n - 1 else mask(N, 4, 0, kmin[i], kmax[i], bw, L)
        else:
            sel, _ = mask(N, 0, k[i], 0, 0, bw, L)
        f = sel.astype(np.complex128)
        I = np.fft.ifftn(f) * N ** 3
        D = np.fft.ifftn(f * spec) * N ** 3
        accI = accI * I
        accD = accD * D
    sD, sI = accD.sum(), accI.sum()
    volume = L * L * L
    val = math.pow(volume, n - 1) * math.pow(N, -n * 3) * sD.real / sI.real
    return val, sI.real / N ** 3, kmin, kmax, int(sel.sum())


def lognormal(n, seed, sigma=0.8, alpha=1.5):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    m = np.fft.fftfreq(n, 1.0 / n)
    m2 = m[:, None, None] ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2
    with np.errstate(divide="ignore"):
        a = np.where(m2 > 0, m2 ** (-alpha / 4), 0.0)
    g = np.fft.ifftn(np.fft.fftn(w) * a).real
    g = (g - g.mean()) / g.std()
    return (np.exp(sigma * g - sigma * sigma / 2) - 1).astype(np.float32)


def main():
    cases = []
    N, L = 16, 100.0
    kf = 2 * PI / L
    field = lognormal(N, 12345)
    spec = np.fft.fftn(field.astype(np.float64))
    configs = []
    for case in (0, 1, 2, 3):
        for (m1, m2, th) in ((3.0, 3.0, 2.0), (2.5, 4.0, 1.1), (4.0, 4.0, 0.0)):
            k1, k2 = m1 * kf, m2 * kf
            k3 = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * math.cos(th))
            lo = [k1, k2, 0.0]
            hi = [k1, k2, 0.0]
            if case in (2, 3):
                lo[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * min(1., math.cos(th) + 0.025))
                hi[2] = math.sqrt(k1 * k1 + k2 * k2 - 2. * k1 * k2 * max(-1., math.cos(th) - 0.025))
            configs.append((3, [k1, k2, k3], lo, hi, case, 1.0))
    for m in (1.0, 3.0, 5.5):
        configs.append((2, [m * kf, m * kf, 0.0], [m * kf, m * kf, 0.0], [m * kf, m * kf, 0.0], 0, 1.0))
        configs.append((2, [m * kf, m * kf, 0.0], [m * kf, m * kf, 0.0], [m * kf, m * kf, 0.0], 1, 1.0))
    configs.append((3, [3 * kf, 3 * kf, 3 * kf], [3 * kf] * 3, [3 * kf] * 3, 0, 2.0))
    for (n, k, lo, hi, case, bw) in configs:
        val, npoly, klo, khi, last = estimator(spec, n, k, lo, hi, bw, case, L)
        cases.append({"n": n, "k": k, "kmin": lo, "kmax": hi, "case": case, "bw": bw,
                      "value": val, "npolygons": npoly,
                      "kmin_out": [float(v).hex() for v in klo], "kmax_out": [float(v).hex() for v in khi],
                      "last_leg_cells": last})
    # mask popcounts and derived extremes on a 32^3 grid
    masks = []
    for (kind, k, lo, hi, bw) in ((0, 7.3, 0, 0, 1.0), (0, 4.5, 0, 0, 1.0), (1, 6.0, 0, 0, 1.0),
                                 (1, 5.0, 0, 0, 1.7), (2, 0, 3.2, 6.1, 1.0), (3, 5.0, 0, 0, 1.0),
                                 (4, 0, 3.0, 4.4, 1.0), (3, 9.9, 0, 0, 2.0)):
        sel, ext = mask(32, kind, k * 2 * PI / 50.0, lo * 2 * PI / 50.0, hi * 2 * PI / 50.0, bw, 50.0)
        ent = {"N": 32, "L": 50.0, "kind": kind, "k": k * 2 * PI / 50.0, "kmin": lo * 2 * PI / 50.0,
               "kmax": hi * 2 * PI / 50.0, "bw": bw, "count": int(sel.sum()),
               "index_sum": int(np.flatnonzero(sel.ravel()).sum())}
        if ext:
            ent["kmin_out"], ent["kmax_out"] = float(ext[0]).hex(), float(ext[1]).hex()
        masks.append(ent)
    out = {"N": N, "L": L, "field": [float(v) for v in field.ravel()], "cases": cases, "masks": masks,
           "generator": "tests/golden/make_golden.py (numpy restatement, numpy.fft)"}
    with open(os.path.join(HERE, "estimator_small.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", len(cases), "cases,", len(masks), "masks")


if __name__ == "__main__":
    main()
