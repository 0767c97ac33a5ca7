"""Summaries of rocprofv3 runs of bench.py (the 512^3 bispectrum sweep).

The committed summaries of the kernel trace, the HBM traffic of a step and
the SQ counters of its largest kernels are written from the CSV files of
rocprofv3 runs of bench.py on an MI355X, by this module, so that every number
in them can be regenerated from the sources.  On the GPU machine, with the
repository at $R and the output at $OUT:

    cd /tmp && export TMPDIR=/tmp
    B="$R/bench.py --steps 20 --warmup 3 --skip-cpu --skip-extras --plain-samples 1 --no-autotune --skip-unrefined"
    rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/trace -o run -- python3 $B
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/fetch -o run -- python3 $B
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/write -o run -- python3 $B
    rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/cfetch -o run -- python3 $R/bench.py --copy-calibration
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/cwrite -o run -- python3 $R/bench.py --copy-calibration
    rocprofv3 --pmc <8 SQ counters> --output-format csv -d $OUT/sq1 -o run -- python3 $B     (optional, ...)
    python3 $R/bench.py > $OUT/bench.json                     (the default run: stage times, tuned launch shape)
    (on any machine with hipcc, for the register counts:)
    hipcc -O3 -std=c++17 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops \\
          --cuda-device-only -S $R/polyspectrum_amd/csrc/ps_kernels.hip -o $OUT/ps_kernels.s
    python3 -m polyspectrum_amd.profile_summary --dir $OUT --method direct --out $R/profiles \\
          --isa $OUT/ps_kernels.s

(add ``--method fourier`` to every bench.py line for the Fourier-space
method).  The traffic summary carries the hash of the kernel sources it was
taken of (``kernel_sources_sha256``); bench.py reports ``roofline.traffic``
from it only when the hash matches the sources it runs.
"""
import argparse
import collections
import csv
import glob
import hashlib
import json
import os
import re
import statistics

_HERE = os.path.dirname(os.path.abspath(__file__))

# the passes of a step (ps_batch_stage_name) and the kernels each one launches
PASSES = {
    "direct": [
        ("G1 masked gather + ky inverse", r"k_inv_g1d?<"),
        ("G2 kx inverse (band -> all x)", r"k_inv_g2<|k_inv_g2_dma<"),
        ("Zc fused C2R + product-reduce", r"k_inv_zc_fused<|k_zc_pairs<|k_zc_dma<"),
        ("final reduce", r"k_reduce_partials"),
    ],
    "fourier": [
        ("G1 masked gather + ky inverse (prefix legs)", r"k_inv_g1d?<"),
        ("G2 kx inverse (band -> all x)", r"k_inv_g2<"),
        ("Zf C2R + product + forward z", r"k_spec_zf<"),
        ("Fx forward x (all x -> band)", r"k_spec_fx<"),
        ("dot forward y + shell sums", r"k_spec_dot<"),
        ("final reduce", r"k_reduce_partials"),
    ],
}
COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_BYTES = 1 << 30  # bench.py --copy-calibration copies 1 GiB per launch


def kernel_sources_sha256(root=None):
    """Hash of the kernel sources (every file of csrc/, by name and content)."""
    csrc = os.path.join(root or _HERE, "csrc")
    h = hashlib.sha256()
    for f in sorted(glob.glob(os.path.join(csrc, "*"))):
        h.update(os.path.basename(f).encode())
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def short(name):
    """Kernel name without its argument list and return type."""
    name = re.sub(r"^void ", "", name)
    depth, out = 0, []
    for ch in name:  # cut at the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    return "".join(out)


def read_trace(d):
    """{kernel: [(dispatch id, start ns, end ns, row)]} of a kernel-trace run."""
    out = collections.defaultdict(list)
    with open(os.path.join(d, "run_kernel_trace.csv")) as f:
        for r in csv.DictReader(f):
            out[short(r["Kernel_Name"])].append((int(r["Dispatch_Id"]), int(r["Start_Timestamp"]),
                                                 int(r["End_Timestamp"]), r))
    for v in out.values():
        v.sort()
    return out


def read_counters(d):
    """{kernel: {dispatch id: {counter: value}}} of a --pmc run."""
    out = collections.defaultdict(lambda: collections.defaultdict(dict))
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for r in csv.DictReader(f):
            out[short(r["Kernel_Name"])][int(r["Dispatch_Id"])][r["Counter_Name"]] = float(r["Counter_Value"])
    return out


def step_kernels(names, method):
    """The FP32 kernels of the mixed-precision step, by pass: {kernel: pass index}
    (the FP64 kernels are the planning run of the polygon counts and bench.py's
    accuracy reference; k_reduce_partials is shared and handled by dispatch range)."""
    sel = {}
    for k in names:
        for i, (_, pat) in enumerate(PASSES[method]):
            if re.search(pat, k) and "<float" in k:
                sel[k] = i
    return sel


def executions(counts):
    """Executions of the step: the most common dispatch count of its kernels."""
    return collections.Counter(counts).most_common(1)[0][0]


def _grid(k):
    m = re.search(r"<float, (\d+)", k)
    return int(m.group(1)) if m else 0


def summarize_trace(tr, method, reps):
    """Per-kernel statistics, the step kernels' pass-by-pass times (the last
    `reps` dispatches are Batch.profile's, one launch at a time) and pass sums."""
    total = sum(e - s for v in tr.values() for _, s, e, _ in v)
    stats = []
    for k, v in tr.items():
        durs = [(e - s) / 1e3 for _, s, e, _ in v]
        stats.append((k, len(v), statistics.mean(durs), sum(durs) / 1e3, 100.0 * sum(durs) * 1e3 / total))
    stats.sort(key=lambda t: -t[3])
    sel = step_kernels(tr.keys(), method)
    nexec = executions([len(tr[k]) for k in sel])
    rows, sums = [], collections.defaultdict(float)
    win_lo, win_hi = None, None
    for k, p in sel.items():
        last = tr[k][-reps:]
        us = statistics.mean((e - s) / 1e3 for _, s, e, _ in last)
        r = last[-1][3]
        rows.append((p, _grid(k), k, len(tr[k]), us, r))
        sums[p] += us
        lo, hi = last[0][1], last[-1][2]
        win_lo = lo if win_lo is None else min(win_lo, lo)
        win_hi = hi if win_hi is None else max(win_hi, hi)
    red = [x for x in tr.get("ps::k_reduce_partials", []) if win_lo <= x[1] and x[2] <= win_hi]
    if red:
        p = len(PASSES[method]) - 1
        us = statistics.mean((e - s) / 1e3 for _, s, e, _ in red)
        rows.append((p, 0, "ps::k_reduce_partials", len(tr["ps::k_reduce_partials"]), us, red[-1][3]))
        sums[p] += us
    rows.sort(key=lambda t: (t[0], t[1], t[2]))
    return stats, rows, sums, nexec


def per_execution(c, ctr, method):
    """{step kernel: counter KiB per execution} of one --pmc run, and the
    number of executions (k_reduce_partials: its dispatches among the step's)."""
    sel = step_kernels(c.keys(), method)
    nexec = executions([len(c[k]) for k in sel])
    ids = [i for k in sel for i in c[k]]
    lo, hi = min(ids), max(ids)
    out = {}
    for k in list(sel) + ["ps::k_reduce_partials"]:
        out[k] = sum(x.get(ctr, 0.0) for i, x in c.get(k, {}).items() if lo <= i <= hi) / nexec
    return out, nexec


def summarize_traffic(fetch, write, cfetch, cwrite, method):
    """Scaled FETCH_SIZE/WRITE_SIZE bytes per execution of each step kernel and per pass."""
    def copy_kib(c, ctr):
        v = [x[ctr] for x in c.get(COPY_KERNEL, {}).values() if ctr in x]
        if not v:
            raise SystemExit(f"no {COPY_KERNEL} dispatch with {ctr} in the calibration run")
        return statistics.median(v)
    ff = COPY_BYTES / 1024 / copy_kib(cfetch, "FETCH_SIZE")
    fw = COPY_BYTES / 1024 / copy_kib(cwrite, "WRITE_SIZE")
    fk, nf = per_execution(fetch, "FETCH_SIZE", method)
    wk, nw = per_execution(write, "WRITE_SIZE", method)
    if nf != nw:
        raise SystemExit(f"the FETCH run executed the step {nf} times, the WRITE run {nw}")
    npass = len(PASSES[method])
    sel = step_kernels(fk.keys(), method)
    rows, per_pass = [], collections.defaultdict(float)
    for k in fk:
        p = sel.get(k, npass - 1)
        gb = (fk[k] * ff + wk.get(k, 0.0) * fw) * 1024 / 1e9
        rows.append((p, _grid(k), k, fk[k], wk.get(k, 0.0), gb))
        per_pass[p] += gb
    rows.sort(key=lambda t: (t[0], t[1], t[2]))
    return rows, per_pass, ff, fw, nf


SQ_DERIVED = [
    ("wait/wave-cycle", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
    ("waiting on an instruction/wave-cycle", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
    ("waiting on LDS/wave-cycle", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
    ("LDS bank-conflict cycles/LDS instruction", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_LDS"),
    ("VALU busy/wave-cycle", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
    ("LDS busy/wave-cycle", "SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
    ("VMEM busy/wave-cycle", "SQ_ACTIVE_INST_VMEM", "SQ_WAVE_CYCLES"),
    ("VALU instructions/LDS instruction", "SQ_INSTS_VALU", "SQ_INSTS_LDS"),
]


def summarize_sq(runs, method, min_grid):
    """Per-dispatch medians of the SQ counters of the step kernels on grids >= min_grid."""
    vals = collections.defaultdict(dict)
    for c in runs:
        sel = step_kernels(c.keys(), method)
        for k in sel:
            if _grid(k) < min_grid:
                continue
            ctrs = collections.defaultdict(list)
            for x in c[k].values():
                for n, v in x.items():
                    ctrs[n].append(v)
            for n, v in ctrs.items():
                vals[k][n] = statistics.median(v)
    return vals


def read_isa(paths):
    """{kernel: (VGPRs, scratch bytes per lane)} from device assembly
    (hipcc --cuda-device-only -S) of the kernel sources."""
    import subprocess
    raw = {}
    for path in paths:
        text = open(path).read()
        for m in re.finditer(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", text, re.S):
            g = lambda key: int(re.search(r"\.amdhsa_" + key + r" (\d+)", m.group(2)).group(1))
            raw[m.group(1)] = (g("next_free_vgpr"), g("private_segment_fixed_size"))
    if not raw:
        return {}
    names = list(raw)
    dem = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True,
                         check=True).stdout.splitlines()
    return {short(d): raw[n] for n, d in zip(names, dem)}


def waves_per_simd(vgpr):
    """Waves one SIMD holds at this VGPR count (512 VGPRs per lane, granule 8, at most 8)."""
    return min(8, 512 // (-(-max(vgpr, 1) // 8) * 8))


def _table(head, rows):
    out = ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)]
    out += ["| " + " | ".join(str(c) for c in r) + " |" for r in rows]
    return "\n".join(out)


def write_summaries(args):
    method = args.method
    names = [p for p, _ in PASSES[method]]
    tag = "bispectrum512" if method == "direct" else "bispectrum512_fourier"
    bench = None
    if args.bench and os.path.exists(args.bench):
        lines = [l for l in open(args.bench).read().splitlines() if l.startswith("{")]
        bench = json.loads(lines[-1]) if lines else None
    cmd_b = ("python3 $R/bench.py --steps 20 --warmup 3 --skip-cpu --skip-extras --plain-samples 1 --no-autotune"
             + (" --method fourier" if method == "fourier" else " --skip-unrefined"))
    regen = (f"python3 -m polyspectrum_amd.profile_summary --dir $OUT --method {method} --out $R/profiles"
             + (" --isa " + " ".join("$OUT/" + os.path.basename(x) for x in args.isa) if args.isa else ""))
    isa_cmd = ("hipcc -O3 -std=c++17 --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops "
               "--cuda-device-only -S $R/polyspectrum_amd/csrc/ps_kernels.hip -o $OUT/ps_kernels.s")
    workload = ("512^3 float32 log-normal field, L = 300 Mpc/h, k1 = k2 in {0.1, 0.4, 0.7, 1.0} h/Mpc x 32 theta "
                "bins (GIVEN binning, dcos = 0.05) = 128 triangles per step, mixed precision, "
                + ("direct method (every leg inverse transformed, product and sum)" if method == "direct"
                   else "Fourier-space method (prefix product forward transformed, shell sums)")
                + ", default launch parameters (`--no-autotune`).")
    os.makedirs(args.out, exist_ok=True)
    src_hash = kernel_sources_sha256()

    tr = read_trace(os.path.join(args.dir, "trace"))
    stats, rows, sums, nexec = summarize_trace(tr, method, args.reps)
    md = [f"# Kernel trace: 512^3 bispectrum sweep, {method} method", "", "Workload: " + workload, "",
          "Commands (GPU machine, repository at $R, output at $OUT):", "", "```",
          "cd /tmp && export TMPDIR=/tmp",
          f"rocprofv3 --kernel-trace --stats --output-format csv -d $OUT/trace -o run -- {cmd_b}",
          *([isa_cmd] if args.isa else []), regen, "```", "",
          f"Kernel sources: sha256 `{src_hash}` (`polyspectrum_amd.profile_summary.kernel_sources_sha256`).", "",
          f"The step's kernels ran {nexec} times: graph replays (warm-up and timed steps) and, last, "
          f"{args.reps} executions pass by pass (`Batch.profile`, one launch at a time, as bench.py's `stages` "
          "and `roofline` time them with HIP events).  Each grid wave has its own instance (`<float, M>`, M the "
          "wave's alias-free transform grid); in the graph the waves run side by side on branch streams, so the "
          "per-call averages of the first table include that sharing of the CUs.  Kernels `<double, ...>` are "
          "the FP64 planning runs of the polygon counts, bench.py's FP64 accuracy reference "
          "(`<double, 512>`: its reference-shaped sample and the condition numbers of the sums, one "
          "reference-shaped evaluation per triangle) and, within each step, the FP64 batch of the triangles the "
          "refinement evaluates again (small grids); `k_direct_pairs` is the transform-free check and "
          "`k_fwd_z`/`k_c2c_*` the forward transforms.  The pass sums below are of the FP32 kernels.", ""]
    md.append(_table(["kernel", "calls", "avg us", "total ms", "% of traced GPU time"],
                     [(f"`{k}`", n, f"{a:.1f}", f"{t:.3f}", f"{p:.2f}") for k, n, a, t, p in stats[:args.top]]))
    isa = read_isa(args.isa) if args.isa else {}
    md += ["", f"Pass by pass (mean of the last {args.reps} dispatches of each step kernel), with the resources "
           "of each kernel: LDS bytes per workgroup and workgroup size (trace), VGPRs per lane and scratch bytes "
           "per lane (code object: `.amdhsa_next_free_vgpr`, `.amdhsa_private_segment_fixed_size` of the device "
           "assembly) and the waves per SIMD those VGPRs allow:", ""]
    md.append(_table(["pass", "kernel", "avg us", "VGPR", "waves/SIMD", "scratch B", "LDS B", "WG"],
                     [(names[p], f"`{k}`", f"{us:.1f}", isa[k][0] if k in isa else "",
                       waves_per_simd(isa[k][0]) if k in isa else "", isa[k][1] if k in isa else "",
                       r["LDS_Block_Size"], r["Workgroup_Size_X"]) for p, _, k, _, us, r in rows]))
    md += ["", "Pass sums from the trace: " + ", ".join(f"{names[p]} {sums[p] / 1e3:.3f} ms" for p in sorted(sums))
           + "."]
    if bench:
        st = bench.get("stages", {})
        md += ["", "The default bench.py run (its own process, autotuned launch parameters "
               f"{bench.get('launch_params')}) timed the same passes with HIP events: "
               + ", ".join(f"{n} {v['ms']:.3f} ms" for n, v in st.items())
               + f"; step (graph replay, results and the FP64 refinement of the flagged triangles, timed on "
               f"the host) {bench.get('ms_per_step')} ms ({bench.get('value')} triangles/s)"
               + (f", {bench['unrefined_step']['ms_per_step']} ms without refinement"
                  if isinstance(bench.get("unrefined_step"), dict) and "ms_per_step" in bench["unrefined_step"] else "")
               + f"; forward transform {bench['forward_fft']['ms']} ms."]
    with open(os.path.join(args.out, f"{tag}_kernel_trace.md"), "w") as f:
        f.write("\n".join(md) + "\n")

    if all(os.path.isdir(os.path.join(args.dir, d)) for d in ("fetch", "write", "cfetch", "cwrite")):
        fetch = read_counters(os.path.join(args.dir, "fetch"))
        write = read_counters(os.path.join(args.dir, "write"))
        cf = read_counters(os.path.join(args.dir, "cfetch"))
        cw = read_counters(os.path.join(args.dir, "cwrite"))
        trows, per_pass, ff, fw, nex = summarize_traffic(fetch, write, cf, cw, method)
        avg = {k: us for _, _, k, _, us, _ in rows}
        js = {"note": f"HBM bytes per execution of the 512^3 sweep (mixed precision, {method} method): "
                      f"FETCH_SIZE x {ff:.3f} + WRITE_SIZE x {fw:.3f} (1 GiB copy calibration), summed over "
                      "each pass's launches (one per grid wave)",
              "kernel_sources_sha256": src_hash,
              "pass_traffic_GB": {names[p]: round(per_pass[p], 5) for p in sorted(per_pass)}}
        with open(os.path.join(args.out, f"{tag}_hbm_traffic.json"), "w") as f:
            json.dump(js, f, indent=1)
            f.write("\n")
        md = [f"# HBM traffic per step: 512^3 bispectrum sweep, {method} method (FETCH_SIZE / WRITE_SIZE)", "",
              "Workload: " + workload, "", "Commands (GPU machine, repository at $R, output at $OUT; one counter per "
              "pass):", "", "```", "cd /tmp && export TMPDIR=/tmp",
              f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/fetch -o run -- {cmd_b}",
              f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/write -o run -- {cmd_b}",
              "rocprofv3 --pmc FETCH_SIZE --output-format csv -d $OUT/cfetch -o run -- python3 $R/bench.py "
              "--copy-calibration",
              "rocprofv3 --pmc WRITE_SIZE --output-format csv -d $OUT/cwrite -o run -- python3 $R/bench.py "
              "--copy-calibration", *([isa_cmd] if args.isa else []), regen, "```", "",
              f"Kernel sources: sha256 `{src_hash}`.", "",
              f"Calibration: a 1 GiB device copy (`{COPY_KERNEL}`) reads {COPY_BYTES / 1024 / ff:.0f} KiB on "
              f"FETCH_SIZE and writes {COPY_BYTES / 1024 / fw:.0f} KiB on WRITE_SIZE, so FETCH_SIZE is scaled by "
              f"{ff:.3f} and WRITE_SIZE by {fw:.3f}.", "",
              f"Per execution of the step ({nex} executions in each run; the counters count L2 misses, i.e. "
              "traffic to the Infinity Cache and HBM): KiB per execution summed over the kernel's dispatches, "
              "and scaled GB.  `avg us` is the pass-by-pass time of the same kernel in the kernel trace "
              f"(`{tag}_kernel_trace.md`).", ""]
        md.append(_table(["pass", "kernel", "FETCH KiB", "WRITE KiB", "traffic GB (scaled)", "avg us"],
                         [(names[p], f"`{k}`", f"{fk:.4g}", f"{wk:.4g}", f"{gb:.4f}",
                           f"{avg[k]:.1f}" if k in avg else "") for p, _, k, fk, wk, gb in trows]))
        line = "Measured per pass: " + ", ".join(f"{names[p]} {per_pass[p]:.3f} GB" for p in sorted(per_pass))
        if bench:
            line += "; algorithmic bytes (bench.py `stages`): " + ", ".join(
                f"{n} {v['GB']:.3f} GB" for n, v in bench.get("stages", {}).items())
        md += ["", line + "."]
        with open(os.path.join(args.out, f"{tag}_hbm_traffic.md"), "w") as f:
            f.write("\n".join(md) + "\n")

    sq_dirs = [os.path.join(args.dir, d) for d in ("sq1", "sq2") if os.path.isdir(os.path.join(args.dir, d))]
    if sq_dirs:
        vals = summarize_sq([read_counters(d) for d in sq_dirs], method, args.sq_min_grid)
        ctrs = sorted({n for v in vals.values() for n in v})
        keys = sorted(vals, key=lambda k: (step_kernels([k], method)[k], _grid(k)))
        res = {k: r for p, _, k, _, us, r in rows}
        md = [f"# SQ counters of the step's kernels on grids >= {args.sq_min_grid}: 512^3 sweep, {method} method",
              "", "Workload: " + workload, "", "Commands (GPU machine, repository at $R, output at $OUT; at most "
              "8 SQ counters per pass):", "", "```", "cd /tmp && export TMPDIR=/tmp"]
        for d in sq_dirs:
            cs = sorted({n for v in read_counters(d).values() for x in v.values() for n in x})
            md.append(f"rocprofv3 --pmc {' '.join(cs)} --output-format csv -d $OUT/{os.path.basename(d)} -o run -- "
                      f"{cmd_b}")
        md += [*([isa_cmd] if args.isa else []), regen, "```", "", f"Kernel sources: sha256 `{src_hash}`.", "",
               "Per-dispatch medians (counter runs serialise the dispatches, so each kernel has the GPU to "
               "itself).  Ratios: a wave-cycle is a cycle of one resident wave (SQ_WAVE_CYCLES), so `x/wave-cycle` "
               "is the fraction of a wave's life spent so; bank-conflict cycles are per LDS instruction.", ""]
        isa = read_isa(args.isa) if args.isa else {}
        head = ["kernel", "VGPR", "LDS B"] + [d for d, a, b in SQ_DERIVED if a in ctrs and b in ctrs]
        trs = []
        for k in keys:
            v = vals[k]
            r = res.get(k, {})
            trs.append([f"`{k}`", isa[k][0] if k in isa else "", r.get("LDS_Block_Size", "")]
                       + [f"{v[a] / v[b]:.3f}" if v.get(b) else "" for d, a, b in SQ_DERIVED
                          if a in ctrs and b in ctrs])
        md.append(_table(head, trs))
        md += ["", "Raw medians:", ""]
        md.append(_table(["kernel"] + ctrs, [[f"`{k}`"] + [f"{vals[k].get(c, float('nan')):.4g}" for c in ctrs]
                                             for k in keys]))
        with open(os.path.join(args.out, f"{tag}_sq_counters.md"), "w") as f:
            f.write("\n".join(md) + "\n")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--dir", required=True, help="rocprofv3 output: trace/, fetch/, write/, cfetch/, cwrite/, sq*/")
    ap.add_argument("--method", choices=sorted(PASSES), default="direct")
    ap.add_argument("--out", required=True, help="folder of the summaries")
    ap.add_argument("--bench", default=None, help="JSON line of the default bench.py run (default: DIR/bench.json)")
    ap.add_argument("--reps", type=int, default=3, help="Batch.profile repetitions at the end of the run")
    ap.add_argument("--top", type=int, default=24, help="kernels listed in the trace table")
    ap.add_argument("--sq-min-grid", type=int, default=96)
    ap.add_argument("--isa", nargs="*", default=None,
                    help="device assembly of the kernel sources (hipcc --cuda-device-only -S), for register counts")
    args = ap.parse_args()
    if args.bench is None:
        args.bench = os.path.join(args.dir, "bench.json")
    write_summaries(args)


if __name__ == "__main__":
    main()
