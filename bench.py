#!/usr/bin/env python3
"""Benchmark: bispectrum triangle configurations per second at 512^3.

Workload: a 512^3 float32 log-normal density field (delta = exp(g - 1/2) - 1,
g a seeded Gaussian random field with P(k) ~ k^-1.5, sigma = 1), box 300
Mpc/h; 4 pairs k1 = k2 in {0.1, 0.4, 0.7, 1.0} h/Mpc x 32 theta bins
(theta_i = pi i / 32, kvectors.c:172-182) = 128 triangles, GIVEN_BINNING with
cosThetaBinwidth = 0.05 (run.c:68-81), kbinwidth = 1.

A step is one evaluation of all 128 triangles (B and Npolygons) by the batched
mixed-precision path (FP32 FFTs, FP64 compensated sums) with the reference's
method: every leg's filtered field inverse transformed, the pointwise product
and its sum.  The triangles of each k1 = k2 pair run in classes on the
smallest grids M on which their products do not alias (M > K1 + K2 + K3 in
cells; M = 16 .. 192 here, the sums scaled by (512/M)^3, exact).  The step replays the batch's captured
hipGraph with the field's spectrum already resident in HBM.  Beside it the
script measures:
  * the opt-in Fourier-space method (same sums in another order: the product
    of the k1, k2 legs forward transformed once, each triangle a sum over its
    k3 shell), reported under "fourier_method";
  * the reference-shaped FP64 path (ps_polyspectrum, one triangle per call)
    on a sample of triangles, and the batched FP64 path on all 128 (the
    accuracy reference for the mixed path);
  * the transform-free sums over the closed lattice triangles
    (ps_direct_check) of all 128, an independent reference for both;
  * per-stage HIP-event times of the batch (the roofline of the dominant
    kernel) and of the forward transform;
  * the CPU baseline: the oracle's FP64 restatement of the reference
    algorithm (C2C FFTs, full-grid masks, products and sums) on one host core.

Prints ONE JSON line.  Only one GPU: --gpus > 1 exits with a message.
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
U32 = 2.0 ** -24  # FP32 unit roundoff
METRIC = "bispectrum triangles/sec at 512³ (mixed vs FP64 plain, ≤1e-5 rel err); HBM GB/s"


def triangles(pairs=(0.1, 0.4, 0.7, 1.0), ntheta=32, dcos=0.05):
    """k-lists as run.c builds them for GIVEN binning (kvectors.c:184-247)."""
    polys = []
    for k in pairs:
        for i in range(ntheta):
            th = math.pi * i / ntheta
            k3 = math.sqrt(k * k + k * k - 2. * k * k * math.cos(th))
            lo = math.sqrt(k * k + k * k - 2. * k * k * min(1., math.cos(th) + 0.5 * dcos))
            hi = math.sqrt(k * k + k * k - 2. * k * k * max(-1., math.cos(th) - 0.5 * dcos))
            polys.append({"n": 3, "k": [k, k, k3], "kmin": [k, k, lo], "kmax": [k, k, hi]})
    return polys


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(N, L):
    """Time the oracle (FP64 restatement of polyspectrum.c) on one core: one
    backward C2C FFT, one mask, one complex product and one complex sum of a
    512^3 grid.  Per triangle (B and Npolygons, as the GPU step returns them)
    the reference runs polyspectrum(): 6 FFTs, 3 masks, 9 products and 6 sums
    (polyspectrum.c:51-161, incl. the 4 diagnostic sums of :142), and
    num_polygons(): 3 FFTs, 3 masks, 3 products and 1 sum (polyspectrum.c:163-252)."""
    import oracle as O
    lib = O.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    n3 = N ** 3
    a = np.empty(2 * n3, dtype=np.float64)
    a[0::2] = np.random.default_rng(1).standard_normal(n3)
    a[1::2] = 0.0
    b = a.copy()
    mask = np.empty(n3, dtype=np.uint8)
    t0 = time.perf_counter()
    lib.or_fft3d(N, a.ctypes.data_as(dp), 1)
    t_fft = time.perf_counter() - t0
    t0 = time.perf_counter()
    lo = ctypes.c_double(0.0)
    hi = ctypes.c_double(0.0)
    lib.or_kfilter(N, 0, 0.7, ctypes.byref(lo), ctypes.byref(hi), 1.0, L, mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    t_mask = time.perf_counter() - t0
    t0 = time.perf_counter()
    lib.or_product(n3, a.ctypes.data_as(dp), b.ctypes.data_as(dp), b.ctypes.data_as(dp))
    t_prod = time.perf_counter() - t0
    s = np.zeros(2)
    t0 = time.perf_counter()
    lib.or_sum(n3, b.ctypes.data_as(dp), s.ctypes.data_as(dp))
    t_sum = time.perf_counter() - t0
    per_tri = 9 * t_fft + 6 * t_mask + 12 * t_prod + 7 * t_sum
    return {
        "value": 1.0 / per_tri,
        "unit": "triangles/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle/ps_oracle.c on 1 core of '{cpu_model()}' (nproc={os.cpu_count()}): one {N}^3 FP64 C2C "
                   f"backward FFT {t_fft:.2f} s, mask {t_mask:.2f} s, complex product {t_prod:.2f} s, complex sum "
                   f"{t_sum:.2f} s, composed as 9 FFT + 6 mask + 12 product + 7 sum = {per_tri:.1f} s per triangle "
                   f"(polyspectrum() + num_polygons(), polyspectrum.c:51-252)"),
    }


def fwd64_refine_ms(P, N, L, field):
    """Forward transform time of the FP64 spectrum a refining context keeps."""
    with P.Context(N, L, P.PREC_FP64) as c:
        c.load_field(field)
        return c.forward_profile(reps=3)[0]


def realisation_batch(P, nper=16, sizes=(128, 192, 256)):
    """Uneven batch of realisations: nper seeded log-normal fields (sigma 0.8) at
    each size, boxes of 1 Mpc/h cells; per field the power spectrum in 32 bins
    k_i = (i+1) k_F and the phase-only bispectrum at 16 theta bins for
    k1 = k2 = 0.3 h/Mpc, DERIVED binning; every batch in one captured graph."""
    from polyspectrum_amd.synthetic import lognormal_field
    print("[bench] realisation batch", file=sys.stderr, flush=True)
    ctxs, batches, npoly = [], [], 0
    for N in sizes:
        L = float(N)
        kf = 2 * math.pi / L
        ctx = P.Context(N, L, P.PREC_MIXED, nslots=nper)
        for r in range(nper):
            ctx.load_field(lognormal_field(N, 3000 + 100 * N + r, sigma=0.8, alpha=1.5), slot=r)
        pk = [{"n": 2, "slot": r, "k": [(i + 1) * kf] * 2, "kmin": [(i + 1) * kf] * 2, "kmax": [(i + 1) * kf] * 2}
              for r in range(nper) for i in range(32)]
        k = 0.3
        bs = []
        for r in range(nper):
            for i in range(16):
                th = math.pi * i / 16
                bs.append({"n": 3, "slot": r, "k": [k, k, math.sqrt(2 * k * k - 2 * k * k * math.cos(th))],
                           "kmin": [0.0] * 3, "kmax": [0.0] * 3})
        ctx.set_phase_only(0)
        batches.append(ctx.batch(pk, 1.0, P.BINNING["DERIVED_BINNING"]))
        ctx.set_phase_only(1)
        batches.append(ctx.batch(bs, 1.0, P.BINNING["DERIVED_BINNING"]))
        npoly += len(pk) + len(bs)
        ctxs.append(ctx)
    mb = P.MultiBatch(batches)
    mb.execute()
    t0 = time.perf_counter()
    reps = 5
    for _ in range(reps):
        mb.execute()
    ms = 1e3 * (time.perf_counter() - t0) / reps
    out = {"fields": nper * len(sizes), "sizes": list(sizes), "configurations": npoly, "ms_per_replay": round(ms, 3),
           "configurations_per_s": round(npoly / (ms / 1e3), 1), "fields_per_s": round(nper * len(sizes) / (ms / 1e3), 1)}
    mb.close()
    for b in batches:
        b.close()
    for c in ctxs:
        c.close()
    return out


def adjoint_timing(P, N):
    """dL/d delta for L = sum_t (B_t - B*_t)^2 over 16 triangles (k1 = k2 = 0.4
    h/Mpc, 16 theta bins, DEFAULT) of a 512^3 log-normal field (sigma 0.8)."""
    from polyspectrum_amd.synthetic import lognormal_field
    print("[bench] adjoint", file=sys.stderr, flush=True)
    L = 400.0 * N / 512
    k = 0.4
    polys = [{"n": 3, "k": [k, k, math.sqrt(2 * k * k - 2 * k * k * math.cos(math.pi * i / 16))],
              "kmin": [0.0] * 3, "kmax": [0.0] * 3} for i in range(16)]
    with P.Context(N, L, P.PREC_MIXED) as ctx:
        ctx.load_field(lognormal_field(N, 4001, sigma=0.8, alpha=1.5))
        b = ctx.batch(polys, 1.0, 0)
        target, _, _ = b.execute()
        ctx.load_field(lognormal_field(N, 4000, sigma=0.8, alpha=1.5))
        v, _, _ = b.execute()
        w = 2.0 * (v - target)
        t0 = time.perf_counter()
        g = b.adjoint(w)
        ms = 1e3 * (time.perf_counter() - t0)
        # the same gradient into a resident float32 torch tensor (no host copy)
        import torch
        gdev = torch.empty(N ** 3, dtype=torch.float32, device="cuda")
        b.adjoint(w, out=gdev)
        torch.cuda.synchronize()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            b.adjoint(w, out=gdev)
        torch.cuda.synchronize()
        ms_dev = 1e3 * (time.perf_counter() - t0) / reps
        dev_vs_host = float(torch.max(torch.abs(gdev.double().cpu() - torch.from_numpy(g.reshape(-1)))) /
                            np.max(np.abs(g)))
        b.close()
    return {"triangles": 16, "grid": N, "ms_per_gradient_on_device": round(ms_dev, 2),
            "ms_per_gradient_to_host": round(ms, 2),
            "includes": "5 full 512^3 transforms, mask applications and products; to_host adds the float64 "
                        "gradient's allocation and device-to-host copy (1 GiB)",
            "device_vs_host_max_rel": dev_vs_host,
            "grad_rms": float(np.sqrt(np.mean(g ** 2)))}


def kernel_sources_sha256():
    """Hash of the kernel sources: a committed counter summary describes the
    build whose sources it names (polyspectrum_amd.profile_summary writes it)."""
    from polyspectrum_amd.profile_summary import kernel_sources_sha256 as h
    return h()


def rocfft_comparison(P, N, L, field, ours):
    """torch.fft (hipFFT -> rocFFT) on the same N^3 field, FP32 and FP64,
    beside the project's forward R2C (ours[prec] = (ms, algorithmic bytes)):
    rfftn (the forward R2C) and irfftn (C2R), HIP events on torch's stream,
    the fastest of several calls after a warm-up (planning) call."""
    import torch
    print("[bench] rocFFT comparison", file=sys.stderr, flush=True)
    out = {"library": "torch.fft -> hipFFT -> rocFFT " + str(getattr(torch.version, "hip", "")), "grid": N}
    for name, dt in (("fp32", torch.float32), ("fp64", torch.float64)):
        x = torch.from_numpy(field).to(device="cuda", dtype=dt)
        res = {}
        for op in ("rfftn", "irfftn"):
            if op == "rfftn":
                f = lambda: torch.fft.rfftn(x)
            else:
                X = torch.fft.rfftn(x)
                f = lambda: torch.fft.irfftn(X, s=x.shape)
            f()
            torch.cuda.synchronize()
            best = 1e30
            for _ in range(5):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                y = f()
                e1.record()
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
                del y
            res[op + "_ms"] = round(best, 4)
        ms, by = ours[name]
        # rocFFT's own model: its input is the field in this precision
        s_in, cs = (4, 8) if name == "fp32" else (8, 16)
        by_roc = s_in * float(N) ** 3 + 5.0 * cs * (N // 2 + 1) * float(N) ** 2
        res["rocfft_bytes_model_GB"] = round(by_roc / 1e9, 4)
        res["rfftn_GB/s"] = round(by_roc / res["rfftn_ms"] / 1e6, 1)
        res["rfftn_frac"] = round(by_roc / res["rfftn_ms"] / 1e6 / HBM_PEAK_GBS, 4)
        res["project_bytes_model_GB"] = round(by / 1e9, 4)  # float32 field in, both precisions
        res["project_forward_ms"] = round(ms, 4)
        res["project_forward_frac"] = round(by / ms / 1e6 / HBM_PEAK_GBS, 4)
        res["project_speedup_vs_rocfft"] = round(res["rfftn_ms"] / ms, 3)
        out[name] = res
        del x
        torch.cuda.empty_cache()
    return out


def scale_err(v, ref, group=32):
    """Largest |v - ref| over the largest |ref| of its k1 = k2 group (32 theta
    bins): the error on the scale of the group's bispectrum, which stays
    meaningful where B(theta) crosses zero and a relative error does not."""
    out = 0.0
    for g in range(0, len(ref), group):
        sl = slice(g, g + group)
        out = max(out, float(np.max(np.abs(v[sl] - ref[sl])) / np.max(np.abs(ref[sl]))))
    return out


def fourier_method(P, N, L, field, polys, steps, warmup, vals64, autotune, kappa):
    """The same sweep by the Fourier-space method (Context.set_method); its
    errors also against the condition-number bar 2 u32 kappa."""
    print("[bench] Fourier-space method", file=sys.stderr, flush=True)
    ctx = P.Context(N, L, P.PREC_MIXED)
    ctx.load_field(field)
    ctx.set_method(P.METHOD_FOURIER)
    if autotune:
        ctx.autotune(polys, 1.0, P.BINNING["GIVEN_BINNING"], reps=2)
    b = ctx.batch(polys, 1.0, P.BINNING["GIVEN_BINNING"])
    for _ in range(warmup):
        b.execute()
    t0 = time.perf_counter()
    for _ in range(steps):
        v, npol, _ = b.execute()
    el = time.perf_counter() - t0
    ms, by, names = b.profile(reps=3)
    params = b.params
    b.close()
    ctx.close()
    rel = np.abs(v - vals64) / np.abs(vals64)
    return {"triangles_per_s": round(len(polys) * steps / el, 1), "ms_per_step": round(1e3 * el / steps, 4),
            "stages": {names[i]: {"ms": round(float(ms[i]), 4), "GB": round(float(by[i]) / 1e9, 4)}
                       for i in range(len(names))},
            "accuracy_vs_fp64_direct": {"rel_err_max": float(rel.max()), "median": float(np.median(rel)),
                                        "fraction_le_1e-5": float(np.mean(rel <= 1e-5)),
                                        "worst_triangle": int(np.argmax(rel)),
                                        "worst_triangle_kappa": float("%.4g" % kappa[int(np.argmax(rel))]),
                                        "max_rel_err_over_u32_kappa": float(np.max(rel / (U32 * kappa))),
                                        "max_err_over_group_scale": scale_err(v, vals64)},
            "launch_params": params}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--plain-samples", type=int, default=4)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--method", choices=("direct", "fourier"), default="direct",
                    help="method of the measured step (direct: the reference's; fourier: opt-in, same sums)")
    ap.add_argument("--skip-extras", action="store_true",
                    help="skip the realisation-batch and adjoint measurements")
    ap.add_argument("--refine-tol", type=float, default=1e-5,
                    help="mixed precision: triangles whose error estimate exceeds this are evaluated again in FP64 "
                         "inside the step (0: off)")
    ap.add_argument("--skip-unrefined", action="store_true",
                    help="do not time the step without refinement beside the refined one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (scale for the FETCH/WRITE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the estimator is not sharded across GPUs in this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import polyspectrum_amd as P
    from polyspectrum_amd.synthetic import lognormal_field

    N = args.size
    L = 300.0 * N / 512
    polys = triangles()
    ntri = len(polys)
    print(f"[bench] generating {N}^3 log-normal field", file=sys.stderr, flush=True)
    field = lognormal_field(N, 2000, sigma=1.0, alpha=1.5)

    # ---- mixed-precision batched path (the measured value)
    ctx = P.Context(N, L, P.PREC_MIXED)
    refine = args.refine_tol > 0 and args.method == "direct"
    if refine:
        ctx.set_refinement(args.refine_tol)  # before the load: the context keeps an FP64 spectrum too
    ctx.load_field(field)
    if args.method == "fourier":
        ctx.set_method(P.METHOD_FOURIER)
    tuned, tuned_ms = None, None
    fwd_tuned = None
    if not args.no_autotune:
        print("[bench] autotuning the forward transform and the launch shape", file=sys.stderr, flush=True)
        fwd_tuned = ctx.forward_autotune(reps=5)
        tuned, tuned_ms = ctx.autotune(polys, 1.0, P.BINNING["GIVEN_BINNING"], reps=5)
    t0 = time.perf_counter()
    batch = ctx.batch(polys, 1.0, P.BINNING["GIVEN_BINNING"])
    plan_ms = 1e3 * (time.perf_counter() - t0)
    for _ in range(args.warmup):
        batch.execute()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vals32, npol32, _ = batch.execute()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = ntri * args.steps / elapsed
    refined, err_est = batch.refinement()
    unrefined = None
    if refine and not args.skip_unrefined:
        # the same batch without refinement (and without the magnitude sums)
        ctx.set_refinement(0.0)
        for _ in range(args.warmup):
            batch.execute()
        nu = max(1, args.steps // 2)
        t0 = time.perf_counter()
        for _ in range(nu):
            vals_u, _, _ = batch.execute()
        eu = time.perf_counter() - t0
        unrefined = {"ms_per_step": round(1e3 * eu / nu, 3), "triangles_per_s": round(ntri * nu / eu, 2),
                     "steps": nu}
        ctx.set_refinement(args.refine_tol)

    # ---- per-pass HIP-event times (dominant pass) and the forward FFT
    used_params = batch.params
    stage_ms, stage_bytes, names = batch.profile(reps=3)
    dom = int(np.argmax(stage_ms))
    traffic, traffic_note = None, "no counter summary for this workload"
    tfile = os.path.join(ROOT, "profiles", "bispectrum512_hbm_traffic.json" if args.method == "direct"
                         else "bispectrum512_fourier_hbm_traffic.json")
    if N == 512 and os.path.exists(tfile):
        # HBM bytes per step of the dominant pass (all its launches) from the
        # committed FETCH_SIZE/WRITE_SIZE summary, if it was taken of these
        # kernel sources
        tj = json.load(open(tfile))
        if tj.get("kernel_sources_sha256") == kernel_sources_sha256():
            t = tj.get("pass_traffic_GB", {}).get(names[dom])
            traffic = None if t is None else round(t * 1e9)
            traffic_note = ("bytes per step over the pass's launches (rocprofv3 FETCH_SIZE+WRITE_SIZE, copy-kernel "
                            "scaled; " + os.path.relpath(tfile, ROOT) + ", taken of these kernel sources)")
        else:
            traffic_note = "committed counter summary was taken of other kernel sources: not used"
    fwd_ms, fwd_bytes = ctx.forward_profile(reps=3)
    # a new field in a refining context also needs its FP64 spectrum
    fwd_refine_ms = fwd64_refine_ms(P, N, L, field) if refine else 0.0
    stages = {names[i]: {"ms": round(float(stage_ms[i]), 4), "GB": round(float(stage_bytes[i]) / 1e9, 4),
                         "GB/s": round(float(stage_bytes[i] / stage_ms[i] / 1e6), 1) if stage_ms[i] > 0 else None}
              for i in range(len(names))}
    batch.close()
    ctx.close()

    # ---- FP64: batched (accuracy reference) and reference-shaped sample
    print("[bench] FP64 paths", file=sys.stderr, flush=True)
    c64 = P.Context(N, L, P.PREC_FP64)
    c64.load_field(field)
    if not args.no_autotune:
        c64.forward_autotune(reps=3)
    fwd64 = c64.forward_profile(reps=3)
    b64 = c64.batch(polys, 1.0, P.BINNING["GIVEN_BINNING"])
    vals64, npol64, _ = b64.execute()
    b64_ms = b64.last_ms
    b64.close()
    # transform-free sums over the closed lattice triangles (independent check)
    t0 = time.perf_counter()
    valsd, npold, _ = c64.direct_check(polys, 1.0, P.BINNING["GIVEN_BINNING"])
    direct_s = time.perf_counter() - t0
    sample = [polys[i] for i in np.linspace(0, ntri - 1, args.plain_samples).astype(int)]
    plain = []
    t0 = time.perf_counter()
    for p in sample:
        plain.append(c64.polyspectrum(3, p["k"], p["kmin"], p["kmax"], 1.0, P.BINNING["GIVEN_BINNING"])[0])
    plain_s = (time.perf_counter() - t0) / len(sample)
    # condition number of every triangle's sum (reference-shaped FP64 path)
    kappa = c64.condition_numbers(polys, 1.0, P.BINNING["GIVEN_BINNING"])
    c64.close()
    idx = np.linspace(0, ntri - 1, args.plain_samples).astype(int)
    plain_vs_b64 = float(np.max(np.abs(np.array(plain) - vals64[idx]) / np.abs(vals64[idx])))
    rel = np.abs(vals32 - vals64) / np.abs(vals64)
    if unrefined is not None:
        rel_u = np.abs(vals_u - vals64) / np.abs(vals64)
        unrefined.update({"rel_err_max": float(rel_u.max()), "fraction_le_1e-5": float(np.mean(rel_u <= 1e-5)),
                          "note": "the same step without refinement: every triangle from the FP32 path"})

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "triangles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 FFT/products, f64 compensated sums; triangles over the error tolerance re-evaluated in f64",
        "data": "synthetic (seeded log-normal field generated on the host)",
        "config": {"workload": "equilateral-pair bispectrum sweep: 512^3 log-normal DENS field, L=300 Mpc/h, "
                               "k1=k2 in {0.1,0.4,0.7,1.0} h/Mpc x 32 theta, GIVEN_BINNING dcos=0.05, kbinwidth=1",
                   "grid": N, "boxsize": L, "triangles_per_step": ntri,
                   "method": ("direct: every leg inverse transformed, product and sum, per group on its "
                              "alias-free grid") if args.method == "direct" else
                             "fourier: prefix product forward transformed, shell sums per last leg"},
        "roofline": {"bound": "hbm", "kernel": names[dom],
                     "achieved": round(float(stage_bytes[dom] / stage_ms[dom] / 1e6), 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(float(stage_bytes[dom] / stage_ms[dom] / 1e6 / HBM_PEAK_GBS), 4),
                     "traffic": traffic, "traffic_note": traffic_note},
        "step_roofline": {"model_GB": round(float(np.sum(stage_bytes)) / 1e9, 4),
                          "GB/s": round(float(np.sum(stage_bytes)) / (ms_per_step * 1e6), 1),
                          "frac": round(float(np.sum(stage_bytes)) / (ms_per_step * 1e6) / HBM_PEAK_GBS, 4)},
        "stages": stages,
        "forward_fft": {"variant": None if fwd_tuned is None else fwd_tuned[0],
                        "variant_note": "radix plan (variant % 2) and tile width (variant // 2), autotuned",
                        "ms": round(fwd_ms, 4), "GB": round(fwd_bytes / 1e9, 4),
                        "GB/s": round(fwd_bytes / fwd_ms / 1e6, 1),
                        "frac": round(fwd_bytes / fwd_ms / 1e6 / HBM_PEAK_GBS, 4)},
        # a step that also transforms the field (a new field per step)
        "with_forward_fft": {"ms_per_step": round(ms_per_step + fwd_ms + fwd_refine_ms, 3),
                             "triangles_per_s": round(ntri / ((ms_per_step + fwd_ms + fwd_refine_ms) / 1e3), 1),
                             "includes": "the FP32 forward transform" + (" and the FP64 one the refinement reads "
                                                                       "(%.3f ms)" % fwd_refine_ms if refine else "")},
        # planning (ps_batch_create: masks, grids, buffers, the FP64 polygon-count
        # transforms of the bare masks, graph capture) and a cold evaluation:
        # plan + forward transform of the field + one step
        "plan": {"ms": round(plan_ms, 2),
                 "cold_ms": round(plan_ms + fwd_ms + fwd_refine_ms + ms_per_step, 2),
                 "includes": "leg masks and bands, alias-free grids, arena, FP64 I-part transforms "
                             "(polygon counts), hipGraph capture and instantiation"},
        "accuracy": {"mixed_vs_fp64_rel_err_max": float(rel.max()), "median": float(np.median(rel)),
                     "fraction_le_1e-5": float(np.mean(rel <= 1e-5)),
                     # triangles whose estimated mixed-precision error u32 (16 + 32 kappa / sqrt(#polygons))
                     # exceeded refine_tol: evaluated again by the FP64 batched path within the step
                     "refinement": {"tol": args.refine_tol if refine else 0.0,
                                    "refined_triangles": np.flatnonzero(refined).tolist(),
                                    "max_rel_err_over_estimate_unrefined":
                                        float(np.max(rel[~refined] / err_est[~refined])) if refine else None},
                     # per triangle (order: k1 = k2 = 0.1, 0.4, 0.7, 1.0, theta_i = pi i / 32):
                     # relative error vs FP64 and the condition number
                     # kappa = sum|prod D| / |sum prod D| of its sum
                     "per_triangle_rel_err": [float("%.3e" % r) for r in rel],
                     "per_triangle_kappa": [float("%.4g" % k) for k in kappa],
                     "max_rel_err_over_u32_kappa": float(np.max(rel / (U32 * kappa))),
                     "max_err_over_group_scale": scale_err(vals32, vals64),
                     "npolygons_rel_err_max": float(np.max(np.abs(npol32 - npol64) / np.abs(npol64))),
                     "plain_fp64_vs_batched_fp64_max": plain_vs_b64,
                     "vs_direct_lattice_sums": {
                         "fp64_batched_rel_err_max": float(np.max(np.abs(vals64 - valsd) / np.abs(valsd))),
                         "mixed_rel_err_max": float(np.max(np.abs(vals32 - valsd) / np.abs(valsd))),
                         "mixed_max_err_over_group_scale": scale_err(vals32, valsd),
                         "npolygons_rounded_equal": bool(np.array_equal(np.round(npol32), npold)),
                         "seconds": round(direct_s, 3)}},
        "fp64_plain": {"triangles_per_s": round(1.0 / plain_s, 3), "sampled_triangles": len(sample),
                       "speedup_mixed_batched": round(value * plain_s, 1)},
        "fp64_batched": {"triangles_per_s": round(ntri / (b64_ms / 1e3), 1)},
        "unrefined_step": unrefined,
        "launch_params": {"used": used_params, "autotuned": tuned is not None,
                          "autotune_best_ms": tuned_ms},
    }
    if not args.skip_extras:
        out["rocfft"] = rocfft_comparison(P, N, L, field, {"fp32": (fwd_ms, fwd_bytes), "fp64": fwd64})
    if not args.skip_extras and args.method == "direct":
        out["fourier_method"] = fourier_method(P, N, L, field, polys, args.steps, args.warmup, vals64,
                                               not args.no_autotune, kappa)
    if not args.skip_extras:
        out["realisations"] = realisation_batch(P)
        out["adjoint"] = adjoint_timing(P, N)
    if not args.skip_cpu:
        print("[bench] CPU baseline (oracle, 1 core)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(N, L)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
