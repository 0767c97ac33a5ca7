"""Uneven batch of realisations at its own sizes: two seeded log-normal fields
(sigma 0.8, P ~ k^-1.5) at each of 128^3, 192^3 and 256^3, boxes of 1 Mpc/h
cells (L = N); per field the power spectrum in 32 bins k_i = (i+1) k_F and
the phase-only bispectrum (useOnlyPhase = 1, run.c:48-49) at 16 theta bins
for k1 = k2 = 0.3 h/Mpc, DERIVED binning (filter.c:78-114, kvectors.c:200-211);
all six batches replayed from one captured graph (ps_multi).

FP64: sampled configurations against the oracle (1e-10), every configuration
against the transform-free lattice sums (1e-10, exact polygon counts).
Mixed: every configuration against FP64 to the mixed bar (tests/_bars.py).
"""
import math

import numpy as np
import pytest

import oracle as O
from _bars import mixed_within_bar, report
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu
PI = math.acos(-1.0)
SIZES = (128, 192, 256)
NREAL = 2
DERIVED = 1


def _polys(N, L, slot, kind):
    kf = 2 * PI / L
    if kind == "P":
        return [{"n": 2, "slot": slot, "k": [(i + 1) * kf] * 2, "kmin": [(i + 1) * kf] * 2,
                 "kmax": [(i + 1) * kf] * 2} for i in range(32)]
    k = 0.3
    out = []
    for i in range(16):
        th = PI * i / 16
        out.append({"n": 3, "slot": slot, "k": [k, k, math.sqrt(2 * k * k - 2 * k * k * math.cos(th))],
                    "kmin": [0.0] * 3, "kmax": [0.0] * 3})
    return out


def _fields(N):
    return [lognormal_field(N, 3000 + 100 * N + r, sigma=0.8, alpha=1.5) for r in range(NREAL)]


def _run(P, prec, with_checks):
    """Values of the six batches from one graph; with_checks (FP64): the
    transform-free sums and the condition numbers of every configuration."""
    ctxs, batches, meta, checks = [], [], [], []
    for N in SIZES:
        L = float(N)
        ctx = P.Context(N, L, prec, nslots=NREAL)
        fields = _fields(N)
        for r, f in enumerate(fields):
            ctx.load_field(f, slot=r)
        for kind, phase in (("P", 0), ("B", 1)):
            polys = sum([_polys(N, L, r, kind) for r in range(NREAL)], [])
            ctx.set_phase_only(phase)
            batches.append(ctx.batch(polys, 1.0, DERIVED))
            meta.append((N, L, kind, polys))
            if with_checks:
                v, npg, _ = ctx.direct_check(polys, 1.0, DERIVED)
                checks.append((v, npg, ctx.condition_numbers(polys, 1.0, DERIVED)))
        ctxs.append(ctx)
    mb = P.MultiBatch(batches)
    res = mb.execute()
    again = mb.execute()
    for (a, _, _), (b, _, _) in zip(res, again):
        np.testing.assert_array_equal(a, b)  # graph replays agree
    mb.close()
    for b in batches:
        b.close()
    for c in ctxs:
        c.close()
    return res, meta, checks


@pytest.fixture(scope="module")
def fp64(gpu_lib):
    return _run(gpu_lib, gpu_lib.PREC_FP64, True)


def test_config3_fp64_vs_oracle_and_lattice_sums(gpu_lib, fp64):
    res, meta, checks = fp64
    for (v, npg, _), (N, L, kind, polys), (vd, npd, _) in zip(res, meta, checks):
        rel = np.abs(v - vd) / np.abs(vd)
        print(N, kind, "FP64 vs lattice sums: max", rel.max())
        assert np.all(rel <= 1e-10), (N, kind, rel.max())
        np.testing.assert_array_equal(np.round(npg), npd)
    # the oracle on sampled configurations: field 1, P bin 8, theta bin 5
    for N in SIZES:
        L = float(N)
        f = _fields(N)[1]
        spec = O.forward_spectrum(f)
        vP = [r[0] for r, m in zip(res, meta) if m[0] == N and m[2] == "P"][0]
        vB = [r[0] for r, m in zip(res, meta) if m[0] == N and m[2] == "B"][0]
        p = _polys(N, L, 1, "P")[7]
        ref = O.polyspectrum(spec, 2, p["k"], p["kmin"], p["kmax"], 1.0, DERIVED, L)[0]
        assert abs(vP[32 + 7] - ref) <= 1e-10 * abs(ref), (N, vP[32 + 7], ref)
        p = _polys(N, L, 1, "B")[5]
        ref = O.polyspectrum(O.calc_phase(spec), 3, p["k"], p["kmin"], p["kmax"], 1.0, DERIVED, L)[0]
        assert abs(vB[16 + 5] - ref) <= 1e-10 * abs(ref), (N, vB[16 + 5], ref)


def test_config3_mixed_vs_fp64(gpu_lib, fp64):
    P = gpu_lib
    res64, meta, checks = fp64
    res32, _, _ = _run(P, P.PREC_MIXED, False)
    for (v32, n32, _), (v64, n64, _), (N, L, kind, polys), (_, _, kappa) in zip(res32, res64, meta, checks):
        rel = np.abs(v32 - v64) / np.abs(v64)
        report(f"{N}^3 {kind}", rel, kappa)
        np.testing.assert_array_equal(n32, n64)
        assert np.all(mixed_within_bar(rel, kappa)), (N, kind, rel.max())
