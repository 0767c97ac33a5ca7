"""GPU parity of the batched path (ps_batch_*): de-duplicated legs, pruned
first inverse passes, fused product-and-reduce, hipGraph replay.

Bars: FP64 batch <= 1e-10 relative to the oracle; mixed precision (FP32
FFTs, FP64 compensated sums) <= 1e-5 relative to the FP64 result per polygon,
or, where the condition number kappa = sum|prod D| / |sum prod D| is large (B
crossing zero), within two FP32 units of sum|prod D| (tests/_bars.py)."""
import math

import numpy as np
import pytest

import oracle as O
from _bars import mixed_within_bar, report
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu


def _polys_triangles(kf, pairs, thetas, case, dcos=0.05):
    out = []
    for (m1, m2) in pairs:
        k1, k2 = m1 * kf, m2 * kf
        for th in thetas:
            k3 = O.calc_k3(k1, k2, math.cos(th))
            lo = [k1, k2, 0.0]
            hi = [k1, k2, 0.0]
            if case in (2, 3):
                lo[2] = O.calc_k3min(k1, k2, th, dcos)
                hi[2] = O.calc_k3max(k1, k2, th, dcos)
            out.append({"n": 3, "k": [k1, k2, k3], "kmin": lo, "kmax": hi})
    return out


def _oracle_values(spec, polys, bw, case, L):
    vals, counts = [], []
    for p in polys:
        v, lo, hi, sums = O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], bw, case, L)
        vals.append(v)
        counts.append(sums[1].real / spec.shape[0] ** 3)
    return np.array(vals), np.array(counts)


# wider bins (bw = 1.5, 2) widen the shells and their radius bounds (band,
# disk and alias-free grid), including the cube-corner WATKINSON tests
@pytest.mark.parametrize("N,L,case,bw", [(32, 100.0, 0, 1.0), (32, 100.0, 2, 1.0), (48, 120.0, 1, 1.0),
                                         (64, 100.0, 3, 1.0), (64, 100.0, 3, 2.0), (48, 120.0, 2, 1.5),
                                         (32, 100.0, 1, 2.0)])
def test_batch_fp64_matches_oracle(gpu_lib, N, L, case, bw):
    P = gpu_lib
    f = lognormal_field(N, 2000 + N)
    spec = O.forward_spectrum(f)
    kf = 2 * math.pi / L
    thetas = [math.pi * i / 6 for i in range(6)]
    polys = _polys_triangles(kf, [(4.0, 4.0), (3.0, 5.0)], thetas, case)
    ref, cnt = _oracle_values(spec, polys, bw, case, L)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        b = ctx.batch(polys, bw, case)
        v, npg, sums = b.execute()
        v2, _, _ = b.execute()  # graph replay
        b.close()
    ok = np.isfinite(ref)
    assert ok.sum() >= len(ref) - 1
    np.testing.assert_allclose(v[ok], ref[ok], rtol=1e-10, atol=0)
    np.testing.assert_array_equal(v, v2)
    np.testing.assert_allclose(npg, cnt, rtol=1e-9, atol=1e-6)
    assert np.all(np.abs(npg - np.round(npg)) < 1e-6)


def test_batch_mixed_vs_fp64(gpu_lib):
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 31)
    kf = 2 * math.pi / L
    thetas = [math.pi * i / 10 for i in range(10)]
    polys = _polys_triangles(kf, [(7.96, 7.96), (5.0, 9.0)], thetas, 2)
    polys += [{"n": 2, "k": [m * kf, m * kf], "kmin": [m * kf, m * kf], "kmax": [m * kf, m * kf]}
              for m in range(1, 20)]
    res = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            b = ctx.batch(polys, 1.0, 2)
            res[prec] = b.execute()
            b.close()
            if prec == P.PREC_FP64:
                kappa = ctx.condition_numbers(polys, 1.0, 2)
    v64, n64, s64 = res[P.PREC_FP64]
    v32, n32, s32 = res[P.PREC_MIXED]
    rel = np.abs(v32 - v64) / np.abs(v64)
    report("64^3 GIVEN", rel, kappa)
    np.testing.assert_array_equal(np.round(n32), np.round(n64))
    assert np.all(mixed_within_bar(rel, kappa)), list(zip(rel[rel > 1e-5], kappa[rel > 1e-5]))


def test_batch_matches_plain_path(gpu_lib):
    """Same configurations through the reference-shaped and the batched path."""
    P = gpu_lib
    N, L = 48, 150.0
    f = lognormal_field(N, 5)
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(6.0, 6.0)], [0.3, 1.5, 2.8], 0)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        plain = [ctx.polyspectrum(3, p["k"], p["kmin"], p["kmax"], 1.0, 0)[0] for p in polys]
        b = ctx.batch(polys, 1.0, 0)
        v, _, _ = b.execute()
        b.close()
    np.testing.assert_allclose(v, plain, rtol=1e-11)


def test_autotune_and_params(gpu_lib):
    """Tuned launch shapes change only summation order, not results; the
    Fourier method gives the direct method's values."""
    P = gpu_lib
    N, L = 32, 100.0
    f = lognormal_field(N, 77)
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(4.0, 4.0)], [0.2, 1.0, 2.0], 0)
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        assert ctx.method == P.METHOD_DIRECT
        b0 = ctx.batch(polys, 1.0, 0, params={"grid_cap": 64, "chunk_tiles": 1, "g2_div": 2})
        v0, _, _ = b0.execute()
        assert b0.params == {"grid_cap": 64, "chunk_tiles": 1, "g2_div": 2, "overlap": 1,
                             "method": P.METHOD_DIRECT}
        b0.close()
        best, ms = ctx.autotune(polys, 1.0, 0, reps=1)
        assert ms > 0 and best["grid_cap"] > 0 and best["method"] == P.METHOD_DIRECT
        b1 = ctx.batch(polys, 1.0, 0)
        assert b1.params == best
        v1, _, _ = b1.execute()
        b1.close()
        ctx.set_method(P.METHOD_FOURIER)
        bf = ctx.batch(polys, 1.0, 0)
        assert bf.params["method"] == P.METHOD_FOURIER
        vf, _, _ = bf.execute()
        bf.close()
    np.testing.assert_allclose(v1, v0, rtol=1e-13)
    np.testing.assert_allclose(vf, v0, rtol=1e-11)


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_MIXED"])
def test_overlap_schedule_matches(gpu_lib, prec_name):
    """Several polygon groups on different transform grids (three triangle
    prefixes, power spectra): the waves run side by side on up to 4, 2 or 3
    branch streams or one after another, with the same values bit for bit, in
    both methods."""
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 41)
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(7.96, 7.96), (5.0, 9.0), (3.0, 3.0)], [0.4, 1.3, 2.2], 2)
    polys += [{"n": 2, "k": [m * kf] * 2, "kmin": [m * kf] * 2, "kmax": [m * kf] * 2} for m in range(1, 12)]
    for method in (P.METHOD_DIRECT, P.METHOD_FOURIER):
        out = {}
        with P.Context(N, L, getattr(P, prec_name)) as ctx:
            ctx.load_field(f)
            for ov in (0, 1, 2, 3):
                b = ctx.batch(polys, 1.0, 2, params={"grid_cap": 512, "chunk_tiles": 16 if method else 2,
                                                     "g2_div": 1, "overlap": ov, "method": method})
                out[ov] = b.execute()
                again = b.execute()
                b.close()
                np.testing.assert_array_equal(out[ov][0], again[0])
        for ov in (1, 2, 3):
            np.testing.assert_array_equal(out[ov][0], out[0][0])
            np.testing.assert_array_equal(out[ov][1], out[0][1])


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_MIXED"])
def test_direct_waves_in_one_arena_region(gpu_lib, prec_name):
    """Direct method with the arena budget below one wave: every wave (grid)
    reuses one buffer region in turn, for the planning run of the polygon
    counts and for every execution; the values equal the side-by-side plan's."""
    import os
    P = gpu_lib
    N, L = 48, 100.0
    f = lognormal_field(N, 57)
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(2.0, 2.0), (4.0, 4.0), (6.5, 6.5)], [0.3, 1.2, 2.6], 2)
    polys += [{"n": 2, "k": [m * kf] * 2, "kmin": [m * kf] * 2, "kmax": [m * kf] * 2} for m in range(1, 9)]
    out = []
    with P.Context(N, L, getattr(P, prec_name)) as ctx:
        ctx.load_field(f)
        for budget in (None, "1"):
            if budget:
                os.environ["PS_ARENA_MB"] = budget
            try:
                b = ctx.batch(polys, 1.0, 2, params={"grid_cap": 1024, "chunk_tiles": 1, "method": P.METHOD_DIRECT})
                out.append(b.execute())
                again = b.execute()
                np.testing.assert_array_equal(out[-1][0], again[0])
                b.close()
            finally:
                os.environ.pop("PS_ARENA_MB", None)
    np.testing.assert_array_equal(out[1][0], out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])
    if prec_name == "PREC_FP64":
        spec = O.forward_spectrum(f)
        for i in (0, 4, 8, 10, 14):
            p = polys[i]
            ref = O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], 1.0, 2, L)[0]
            assert abs(out[0][0][i] - ref) <= 1e-10 * abs(ref), (i, out[0][0][i], ref)


def test_direct_groups_split_by_grid(gpu_lib):
    """Triangles sharing k1 = k2 whose third legs need different alias-free
    grids (M > K1 + K2 + K3) run as classes on their own grids, one wave per
    grid; each value equals the triangle evaluated alone and the oracle."""
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 91)
    kf = 2 * math.pi / L
    k = 8.0 * kf
    polys = [{"n": 3, "k": [k, k, m * kf], "kmin": [k, k, m * kf], "kmax": [k, k, m * kf]}
             for m in (1.0, 10.0, 15.0, 16.0)]
    params = {"grid_cap": 1024, "chunk_tiles": 1, "method": P.METHOD_DIRECT}
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        b = ctx.batch(polys, 1.0, 0, params=params)
        grids = [w[0] for w in b.waves()]
        assert len(set(grids)) >= 2 and grids == sorted(grids) and max(grids) <= N, grids
        assert all(w[1] >= 1 and w[2] >= 1 and w[3] > 0 for w in b.waves())
        v, npg, _ = b.execute()
        b.close()
        for i, p in enumerate(polys):
            b1 = ctx.batch([p], 1.0, 0, params=params)
            assert len(b1.waves()) == 1
            v1, npg1, _ = b1.execute()
            b1.close()
            assert abs(v[i] - v1[0]) <= 1e-12 * abs(v1[0]), (i, v[i], v1[0])
            assert npg[i] == npg1[0]
    spec = O.forward_spectrum(f)
    for i, p in enumerate(polys):
        ref = O.polyspectrum(spec, 3, p["k"], p["kmin"], p["kmax"], 1.0, 0, L)[0]
        assert abs(v[i] - ref) <= 1e-10 * abs(ref), (i, v[i], ref)


@pytest.mark.parametrize("N,L,m,grid", [(192, 200.0, (40.0, 40.0, (50.0, 55.0, 70.0)), 144),
                                         (216, 250.0, (60.0, 60.0, (80.0, 85.0, 30.0)), 216)])
def test_reduced_grid_3x3(gpu_lib, N, L, m, grid):
    """Grids 2^a 3^b with b >= 2: triangles whose product needs 128 < M <= 144
    run on the 144 grid (radix 3, 3, 8, 2) inside a 192 context; on a 216
    context (radix 3, 3, 3, 8; wavefront lines and tile columns not multiples
    of 16) triangles run on the full grid or a reduced one.  FP64 equals the
    oracle and the mixed path equals FP64 to the mixed bar."""
    P = gpu_lib
    f = lognormal_field(N, grid)
    kf = 2 * math.pi / L
    polys = [{"n": 3, "k": [m[0] * kf, m[1] * kf, m3 * kf], "kmin": [0.0] * 3, "kmax": [0.0] * 3} for m3 in m[2]]
    res = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            b = ctx.batch(polys, 1.0, 0, params={"grid_cap": 1024, "chunk_tiles": 1, "method": P.METHOD_DIRECT})
            grids = [w[0] for w in b.waves()]
            res[prec] = b.execute()
            b.close()
        assert grid in grids, grids
    spec = O.forward_spectrum(f)
    for i in (0, 2):
        p = polys[i]
        ref = O.polyspectrum(spec, 3, p["k"], p["kmin"], p["kmax"], 1.0, 0, L)[0]
        assert abs(res[P.PREC_FP64][0][i] - ref) <= 1e-10 * abs(ref), (i, res[P.PREC_FP64][0][i], ref)
    rel = np.abs(res[P.PREC_MIXED][0] - res[P.PREC_FP64][0]) / np.abs(res[P.PREC_FP64][0])
    assert np.all(rel < 1e-5), rel


def test_reduced_grid_by_euclidean_radii(gpu_lib):
    """k1 = k2 = 10.2 k_F, k3 = 9.9 k_F: the per-axis bands (11 + 11 + 11 =
    33) would need M = 48, the shells' radii (10.7 + 10.7 + 10.4 = 31.8) allow
    M = 32, where no nonzero closed combination of their cells can alias onto
    m = 0.  FP64 equals the oracle, mixed equals FP64 to the mixed bar, and the
    counts are the oracle's."""
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 41)
    kf = 2 * math.pi / L
    polys = [{"n": 3, "k": [10.2 * kf, 10.2 * kf, 9.9 * kf], "kmin": [0.0] * 3, "kmax": [0.0] * 3},
             {"n": 3, "k": [10.2 * kf, 10.2 * kf, 6.0 * kf], "kmin": [0.0] * 3, "kmax": [0.0] * 3}]
    res = {}
    for prec in (P.PREC_FP64, P.PREC_MIXED):
        with P.Context(N, L, prec) as ctx:
            ctx.load_field(f)
            b = ctx.batch(polys, 1.0, 0, params={"grid_cap": 1024, "chunk_tiles": 1, "method": P.METHOD_DIRECT})
            grids = [w[0] for w in b.waves()]
            res[prec] = b.execute()
            b.close()
        assert grids == [32], grids
    spec = O.forward_spectrum(f)
    for i, p in enumerate(polys):
        ref, _, _, rs = O.polyspectrum(spec, 3, p["k"], p["kmin"], p["kmax"], 1.0, 0, L)
        assert abs(res[P.PREC_FP64][0][i] - ref) <= 1e-10 * abs(ref), (i, res[P.PREC_FP64][0][i], ref)
        assert round(res[P.PREC_FP64][1][i]) == round(rs[1].real / N ** 3) > 0
    rel = np.abs(res[P.PREC_MIXED][0] - res[P.PREC_FP64][0]) / np.abs(res[P.PREC_FP64][0])
    assert np.all(rel < 1e-5), rel


@pytest.mark.parametrize("method", ["METHOD_DIRECT", "METHOD_FOURIER"])
def test_open_polygons_on_reduced_grids(gpu_lib, method):
    """Polygons that cannot close (k3 > k1 + k2; a cross spectrum of two
    separate shells) run on grids smaller than their widest shell (K = 3, 3,
    14 on M = 24): their sums of D and I products are zero, as the oracle's,
    and closed polygons in the same batch keep the oracle's values."""
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 23)
    kf = 2 * math.pi / L
    z3, z2 = [0.0] * 3, [0.0] * 2
    polys = [{"n": 3, "k": [3 * kf, 3 * kf, 14 * kf], "kmin": z3, "kmax": z3},
             {"n": 2, "k": [3 * kf, 14 * kf], "kmin": z2, "kmax": z2},
             {"n": 3, "k": [3 * kf, 3 * kf, 5 * kf], "kmin": z3, "kmax": z3},
             {"n": 2, "k": [3 * kf, 3 * kf], "kmin": z2, "kmax": z2}]
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(f)
        ctx.set_method(getattr(P, method))
        b = ctx.batch(polys, 1.0, 0, params={"grid_cap": 512, "chunk_tiles": 1, "method": getattr(P, method)})
        grids = [w[0] for w in b.waves()]
        v, npg, sums = b.execute()
        b.close()
    assert min(grids) < 2 * 14, grids
    spec = O.forward_spectrum(f)
    ref = [O.polyspectrum(spec, p["n"], p["k"], p["kmin"], p["kmax"], 1.0, 0, L) for p in polys]
    for closed in (2, 3):
        assert abs(v[closed] - ref[closed][0]) <= 1e-10 * abs(ref[closed][0])
    for open_, closed in ((0, 2), (1, 3)):
        sd, si = abs(sums[closed][0]), abs(sums[closed][1])
        assert abs(ref[open_][3][0]) <= 1e-12 * sd and abs(ref[open_][3][1]) <= 1e-12 * si
        assert abs(sums[open_][0]) <= 1e-10 * sd, (open_, sums[open_], sums[closed])
        assert abs(sums[open_][1]) <= 1e-10 * si, (open_, sums[open_], sums[closed])


@pytest.mark.parametrize("prec_name", ["PREC_FP64", "PREC_MIXED"])
def test_persistent_grid_sizes_agree(gpu_lib, prec_name):
    """A persistent grid of 3 workgroups runs every pass through its
    prefetching tile loop many times (next tile loaded while the current one
    is transformed); the values equal those of a grid with one workgroup per
    tile, bit for bit, with either G2 tile width."""
    P = gpu_lib
    N, L = 64, 100.0
    f = lognormal_field(N, 3)
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(7.0, 7.0), (4.0, 9.0)], [0.5, 1.4, 2.5], 2)
    out = {}
    with P.Context(N, L, getattr(P, prec_name)) as ctx:
        ctx.load_field(f)
        for cap in (3, 4096):
            for div in (1, 2):
                b = ctx.batch(polys, 1.0, 2, params={"grid_cap": cap, "chunk_tiles": 1, "g2_div": div,
                                                     "method": P.METHOD_DIRECT})
                out[(cap, div)] = b.execute()
                b.close()
    for div in (1, 2):
        np.testing.assert_array_equal(out[(3, div)][0], out[(4096, div)][0])
        np.testing.assert_array_equal(out[(3, div)][1], out[(4096, div)][1])
    np.testing.assert_allclose(out[(3, 1)][0], out[(3, 2)][0], rtol=1e-12 if prec_name == "PREC_FP64" else 1e-5)


def test_context_close_closes_its_batches(gpu_lib):
    """Closing a context first closes the batches and multi-batches still
    open on it; their later close (or collection) does nothing."""
    P = gpu_lib
    N, L = 32, 100.0
    kf = 2 * math.pi / L
    polys = _polys_triangles(kf, [(3.0, 3.0)], [0.4, 1.3], 2)
    ctx = P.Context(N, L, P.PREC_FP64)
    ctx.load_field(lognormal_field(N, 43))
    b1 = ctx.batch(polys, 1.0, 2)
    b2 = ctx.batch(polys, 1.0, 2)
    v1 = b1.execute()[0]
    m = P.MultiBatch([b1, b2])
    out = m.execute()
    np.testing.assert_array_equal(out[0][0], v1)
    ctx.close()
    assert b1._h is None and b2._h is None and m._h is None
    b1.close()
    m.close()
