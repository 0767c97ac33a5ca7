"""Mixed precision with FP64 refinement of ill-conditioned polygons
(ps_set_refinement, ps_batch_refinement).

A refined polygon's value is the FP64 batched path's: it must equal an FP64
batch of the same polygons to FP64 rounding (1e-12).  Unrefined values are
the mixed path's.  The error estimate u32 (16 + 32 kappa / sqrt(#polygons))
must bound every measured mixed-vs-FP64 error, and with tol = 1e-5 the
benchmark's 512^3 sweep (128 triangles) is within 1e-5 of FP64 everywhere."""
import math

import numpy as np
import pytest

import oracle as O
from polyspectrum_amd.synthetic import lognormal_field

pytestmark = pytest.mark.gpu


def _triangles(kf, pairs, nth, dcos=0.05):
    out = []
    for (m1, m2) in pairs:
        k1, k2 = m1 * kf, m2 * kf
        for i in range(nth):
            th = math.pi * i / nth
            k3 = O.calc_k3(k1, k2, math.cos(th))
            out.append({"n": 3, "k": [k1, k2, k3], "kmin": [k1, k2, O.calc_k3min(k1, k2, th, dcos)],
                        "kmax": [k1, k2, O.calc_k3max(k1, k2, th, dcos)]})
    return out


def _fp64_values(P, N, L, field, polys, case):
    with P.Context(N, L, P.PREC_FP64) as ctx:
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        v, npg, s = b.execute()
        b.close()
    return v, npg, s


def test_refine_everything_equals_fp64(gpu_lib):
    """tol far below any estimate: every polygon comes from the FP64 path."""
    P = gpu_lib
    N, L = 64, 100.0
    field = lognormal_field(N, 7001)
    polys = _triangles(2 * math.pi / L, [(5.0, 5.0), (4.0, 7.0)], 8)
    case = P.BINNING["GIVEN_BINNING"]
    v64, n64, s64 = _fp64_values(P, N, L, field, polys, case)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.set_refinement(1e-300)
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        v, npg, s = b.execute()
        flags, est = b.refinement()
        b.close()
    assert flags.all()
    np.testing.assert_allclose(v, v64, rtol=1e-12, atol=0)
    np.testing.assert_array_equal(npg, n64)
    assert np.all(est > 0)


def test_estimates_bound_errors_and_nothing_refined_at_infinite_tol(gpu_lib):
    P = gpu_lib
    N, L = 128, 100.0
    field = lognormal_field(N, 7002, sigma=1.0, alpha=1.5)
    polys = _triangles(2 * math.pi / L, [(6.0, 6.0), (10.0, 10.0), (5.0, 9.0)], 16)
    case = P.BINNING["GIVEN_BINNING"]
    v64, _, _ = _fp64_values(P, N, L, field, polys, case)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        v_plain, _, _ = b.execute()
        f0, e0 = b.refinement()
        assert not f0.any() and np.all(e0 == 0)  # no refinement: no estimates
        ctx.set_refinement(math.inf)  # estimates only (the batch re-captures its passes)
        v, _, _ = b.execute()
        flags, est = b.refinement()
        b.close()
    assert not flags.any()
    rel = np.abs(v - v64) / np.abs(v64)
    rel_plain = np.abs(v_plain - v64) / np.abs(v64)
    print(f"max rel {rel.max():.3e}, max rel / estimate {np.max(rel / est):.3f}, largest estimate {est.max():.3e}")
    assert np.all(rel <= est), np.max(rel / est)
    assert np.all(rel_plain <= est), np.max(rel_plain / est)


def test_refinement_enabled_after_load_and_batch(gpu_lib):
    """The field still in the upload buffer gets its FP64 spectrum when
    refinement is switched on; an existing batch re-captures its passes."""
    P = gpu_lib
    N, L = 64, 100.0
    field = lognormal_field(N, 7003)
    polys = _triangles(2 * math.pi / L, [(6.0, 6.0)], 6)
    case = P.BINNING["GIVEN_BINNING"]
    v64, _, _ = _fp64_values(P, N, L, field, polys, case)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        b.execute()
        ctx.set_refinement(1e-300)
        v, _, _ = b.execute()
        assert b.refinement()[0].all()
        np.testing.assert_allclose(v, v64, rtol=1e-12, atol=0)
        # a new field reaches the FP64 spectrum too
        f2 = lognormal_field(N, 7004)
        ctx.load_field(f2)
        v2, _, _ = b.execute()
        ctx.set_refinement(0.0)
        v3, _, _ = b.execute()
        assert not b.refinement()[0].any()
        b.close()
    w64, _, _ = _fp64_values(P, N, L, f2, polys, case)
    np.testing.assert_allclose(v2, w64, rtol=1e-12, atol=0)
    assert np.all(np.abs(v3 - w64) <= 1e-3 * np.abs(w64))


def test_refinement_rejected_for_fp64_context(gpu_lib):
    P = gpu_lib
    with P.Context(32, 100.0, P.PREC_FP64) as ctx:
        with pytest.raises(P.GPUError):
            ctx.set_refinement(1e-5)
        ctx.set_refinement(0.0)


def test_config_512_sweep_within_1e5_with_refinement(gpu_lib):
    """The benchmark sweep (128 triangles at 512^3): with tol = 1e-5 every
    value is within 1e-5 of the FP64 batch; the refined ones equal it."""
    P = gpu_lib
    import bench
    N, L = 512, 300.0
    field = lognormal_field(N, 2000, sigma=1.0, alpha=1.5)
    polys = bench.triangles()
    case = P.BINNING["GIVEN_BINNING"]
    v64, _, _ = _fp64_values(P, N, L, field, polys, case)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.set_refinement(1e-5)
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        v, _, _ = b.execute()
        flags, est = b.refinement()
        v_again, _, _ = b.execute()
        b.close()
    rel = np.abs(v - v64) / np.abs(v64)
    print(f"refined {np.flatnonzero(flags).tolist()}, max rel {rel.max():.3e}")
    assert np.all(rel <= 1e-5), rel.max()
    np.testing.assert_allclose(v[flags], v64[flags], rtol=1e-12, atol=0)
    assert np.all(rel[~flags] <= est[~flags])
    np.testing.assert_array_equal(v_again, v)  # replays (and the cached FP64 set) reproduce bit for bit


def test_magnitude_sums_not_below_kappa_numerator(gpu_lib):
    """The refinement's numerator of kappa (sum |prod D| on every x plane of
    grids up to 192) against the exact sum |prod D| of the FP64
    reference-shaped path, on the benchmark sweep: equal up to the
    mixed-precision sums (the measured ratio is 0.997-1.0)."""
    P = gpu_lib
    import bench
    N, L = 512, 300.0
    field = lognormal_field(N, 2000, sigma=1.0, alpha=1.5)
    polys = bench.triangles()
    case = P.BINNING["GIVEN_BINNING"]
    with P.Context(N, L, P.PREC_FP64) as c64:
        c64.load_field(field)
        kappa = c64.condition_numbers(polys, 1.0, case)
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.set_refinement(math.inf)  # estimates only
        ctx.load_field(field)
        b = ctx.batch(polys, 1.0, case)
        _, npol, _ = b.execute()
        _, est = b.refinement()
        b.close()
    u32 = 2.0 ** -24
    kappa_used = (est / u32 - 16.0) * np.sqrt(np.maximum(npol, 1.0)) / 32.0
    ratio = kappa_used / kappa
    print(f"numerator used / exact: min {ratio.min():.3f} median {np.median(ratio):.3f} max {ratio.max():.3f}")
    assert np.all(ratio >= 0.99), (np.argmin(ratio), ratio.min())
    assert np.all(ratio <= 1.01), (np.argmax(ratio), ratio.max())


def test_refined_values_equal_across_schedules(gpu_lib):
    """The refined step (magnitude sums, the FP64 batch beside it) gives the
    same values bit for bit with its waves one after another and on two or
    three branch streams, over repeated executions."""
    P = gpu_lib
    N, L = 128, 100.0
    field = lognormal_field(N, 7005, sigma=1.0, alpha=1.5)
    polys = _triangles(2 * math.pi / L, [(6.0, 6.0), (10.0, 10.0), (5.0, 9.0)], 16)
    case = P.BINNING["GIVEN_BINNING"]
    out = {}
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.set_refinement(1e-6)
        ctx.load_field(field)
        for ov in (0, 2, 3):
            b = ctx.batch(polys, 1.0, case, params={"grid_cap": 1024, "chunk_tiles": 1, "g2_div": 1, "overlap": ov,
                                                    "method": P.METHOD_DIRECT})
            vals = [b.execute()[0] for _ in range(3)]
            flags, _ = b.refinement()
            b.close()
            for v in vals[1:]:
                np.testing.assert_array_equal(v, vals[0])
            out[ov] = (vals[0], flags)
    assert out[0][1].any()  # some polygons are refined at this tolerance
    for ov in (2, 3):
        np.testing.assert_array_equal(out[ov][1], out[0][1])
        np.testing.assert_array_equal(out[ov][0], out[0][0])
