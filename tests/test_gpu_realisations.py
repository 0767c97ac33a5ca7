"""Uneven batches of realisations in one captured graph (ps_multi_*): several
grid sizes, several field slots per size, power spectra (amplitude) and
phase-only bispectra under DERIVED binning, all replayed together.

FP64 contexts: every value against the oracle (1e-10 relative).  Mixed
contexts: against the FP64 values, 1e-5 of the group's largest |value|."""
import math

import numpy as np
import pytest

import oracle as O
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)
SIZES = (32, 48, 64)
NREAL = 2


def _polys(N, L, slot, kind):
    kf = 2 * PI / L
    if kind == "P":
        return [{"n": 2, "slot": slot, "k": [(i + 1) * kf] * 2, "kmin": [(i + 1) * kf] * 2,
                 "kmax": [(i + 1) * kf] * 2} for i in range(8)]
    k = 4.0 * kf
    out = []
    for i in range(16):
        th = PI * i / 16
        k3 = math.sqrt(2 * k * k - 2 * k * k * math.cos(th))
        out.append({"n": 3, "slot": slot, "k": [k, k, k3], "kmin": [k, k, 0.0], "kmax": [k, k, 0.0]})
    return out


def _run(P, prec):
    ctxs, batches, meta = [], [], []
    for N in SIZES:
        L = float(N)  # 1 Mpc/h cells
        ctx = P.Context(N, L, prec, nslots=NREAL)
        fields = [lognormal_field(N, 3000 + 10 * N + r, sigma=0.8, alpha=1.5) for r in range(NREAL)]
        for r, f in enumerate(fields):
            ctx.load_field(f, slot=r)
        ctx.set_phase_only(0)
        bp = ctx.batch(sum([_polys(N, L, r, "P") for r in range(NREAL)], []), 1.0, P.BINNING["DERIVED_BINNING"])
        ctx.set_phase_only(1)
        bb = ctx.batch(sum([_polys(N, L, r, "B") for r in range(NREAL)], []), 1.0, P.BINNING["DERIVED_BINNING"])
        ctxs.append(ctx)
        batches += [bp, bb]
        meta += [(N, L, fields, "P"), (N, L, fields, "B")]
    mb = P.MultiBatch(batches)
    res = mb.execute()
    res2 = mb.execute()
    for (a, _, _), (b, _, _) in zip(res, res2):
        np.testing.assert_array_equal(a, b)  # replays agree
    out = [r[0] for r in res]
    mb.close()
    for b in batches:
        b.close()
    for c in ctxs:
        c.close()
    return out, meta


def test_realisation_batch_fp64_vs_oracle(gpu_lib):
    P = gpu_lib
    vals, meta = _run(P, P.PREC_FP64)
    for v, (N, L, fields, kind) in zip(vals, meta):
        per = len(v) // NREAL
        for r, f in enumerate(fields):
            spec = O.forward_spectrum(f)
            if kind == "B":
                spec = O.calc_phase(spec)
            for j, p in enumerate(_polys(N, L, r, kind)):
                ref = O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], 1.0, 1, L)[0]
                got = v[r * per + j]
                assert abs(got - ref) <= 1e-10 * abs(ref), (N, kind, r, j, got, ref)


def test_realisation_batch_mixed_vs_fp64(gpu_lib):
    P = gpu_lib
    v64, meta = _run(P, P.PREC_FP64)
    v32, _ = _run(P, P.PREC_MIXED)
    for a, b, (N, L, _, kind) in zip(v32, v64, meta):
        err = np.abs(a - b).max() / np.abs(b).max()
        print(N, kind, err)
        assert err < 1e-5
