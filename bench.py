"""Benchmark: tracked pixels per second on the 4096^2, sigma = 15 full-disk
pair (64 x 64 boxes, every pixel tracked), one MI355X.

One step = one flct_hip_track_device call over the resident image pair:
box means, the tracker kernel over all 16,777,216 pixels, and the final
reset.  Inputs are in HBM before the timed region.  Prints one JSON line.
"""
import argparse
import collections
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "tracked px/s at 4096^2 sigma=15: split-BF16 MFMA vs FP64 plain; RMS vel err"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_TFLOPS = 2500.0
PEAK_LDS_TBS = 128 * 256 * 2.4e9 / 1e12  # 128 B per clock per CU, 256 CUs, 2.4 GHz
PMC_FILE = os.path.join(ROOT, "profiles", "track_4096_sigma15_traffic.json")


def w_fft(n):
    """Algorithmic flops per tracked pixel of the FFT form: three real 2-D
    FFTs (2.5 N^2 log2 N^2 each) plus window, mean, filter, product, |.|,
    argmax: 15 N^2 log2 N + 14 N^2."""
    return 15.0 * n * n * math.log2(n) + 14.0 * n * n


def w_gemm(n):
    """Contraction flops per tracked pixel of the split-BF16 form as dense
    DFT-matrix products: forward 2 [4 N^3 + 8 N^2 (N/2+1)] (two real images),
    inverse 8 N^2 (N/2+1) + 4 N^3, times 3 for hi*hi + hi*lo + lo*hi."""
    h = n // 2 + 1
    return 3.0 * (2 * (4 * n ** 3 + 8 * n * n * h) + (8 * n * n * h + 4 * n ** 3))


def lds_bytes_fft(n, elem=16, window_reads=1):
    """LDS bytes per tracked pixel of the FFT tracker k_track (complex
    element of elem bytes, window values double), counted from its access
    pattern: phase 1 stores the N x N packed column spectra and reads the two
    row-window values of every box element; phase 2 reads rows k and N-k,
    parks and rereads A's spectrum and stores the product's inverse
    (5N elements per row pair, 10N for rows 0 and N/2); phase 3 reads the
    N/2+1 stored rows as N columns.  4.5 N^2 + 5N elements plus 2 N^2
    doubles per window pass (the FP32 kernels for boxes 44..96 make two:
    the box maxima, then the scaled values)."""
    return (4.5 * n * n + 5 * n) * elem + window_reads * 2 * n * n * 8


MFMA_FLOPS = 2 * 16 * 16 * 32  # one v_mfma_f32_16x16x32_bf16


def seg_mfma_flops(n):
    """MFMA flops the row-segment split-BF16 kernel executes per interior
    pixel (boxes 36..64): 112 three-product groups per pixel (stage 2: 64,
    inverse 1: 32, inverse 2: 16), plus the two products shared by a 64-pixel
    segment: the box-row transform (128 groups) and the mean-correction
    transforms of its pixels (64 groups)."""
    return (112 + (128 + 64) / 64) * 3 * MFMA_FLOPS


def r2_mfma_flops(n, bias=False):
    """MFMA flops the per-pixel radix-2 kernel executes per tracked pixel
    (clamped edge columns and the bias pass): per wave and pass 88
    three-product groups, two waves per pixel."""
    return 88 * 3 * MFMA_FLOPS * 2 * (2 if bias else 1)


def form_roofline(form, r, nbox):
    """Algorithmic FFT-form flops (w_fft) per tracked pixel over the tracking
    kernel's HIP-event time, against the form's vector peak."""
    peak = PEAK_FP64_TFLOPS if form == "fp64" else PEAK_FP32_TFLOPS
    achieved = w_fft(nbox) * r["tracked"] / (r["main_ms"] * 1e-3) / 1e12
    return {"achieved_TFLOPs": round(achieved, 3), "peak_TFLOPs": peak, "frac": round(achieved / peak, 4),
            "flops_per_px": w_fft(nbox)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """Model name and thread count of this host's CPUs, and whether a host
    FFTW3 library is installed (ldconfig -p; ROCm's GPU libhipfftw does not
    count)."""
    model = "unknown"
    fields = {}
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if ":" not in line:
                    if fields:
                        break  # the first processor's block
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                fields.setdefault(k, v)
    except OSError:
        pass
    if "model name" in fields:
        model = fields["model name"]
        # a virtualised host may report only the brand: add the family,
        # model and stepping numbers, which identify the core generation
        extra = [f"{k} {fields[k]}" for k in ("cpu family", "model", "stepping") if k in fields]
        if "hypervisor" in fields.get("flags", "").split():
            extra.append("virtual machine")
        if extra:
            model = f"{model} ({', '.join(extra)})"
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    fftw = None
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            name = line.strip().split(" ", 1)[0]
            if name.startswith("libfftw3.so"):
                fftw = line.split("=>")[-1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "threads_available": avail,
            "host_libfftw3": fftw}


def cpu_baseline(f1, f2, sigma, budget_s=20.0):
    """The CPU oracle (single-thread C restatement of the reference loop) on a
    band of rows through the disk centre; returns (px/s, sample text)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    n = f1.shape[1]
    c = f1.shape[0] // 2
    probe = np.arange(c * n + n // 2 - 256, c * n + n // 2 + 256)
    t = time.perf_counter()
    oracle_lib.flct_pixels(f1, f2, 1.0, 1.0, sigma, probe)
    per_px = (time.perf_counter() - t) / probe.size
    rows = int(max(1, min(32, round(budget_s / per_px / n))))
    r0 = c - rows // 2
    pix = np.arange(r0 * n, (r0 + rows) * n)
    t = time.perf_counter()
    oracle_lib.flct_pixels(f1, f2, 1.0, 1.0, sigma, pix)
    dt = time.perf_counter() - t
    return pix.size / dt, f"rows {r0}-{r0 + rows - 1} of the 4096^2 sigma=15 pair ({pix.size} px), 1 thread"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--sigma", type=float, default=15.0)
    ap.add_argument("--forms", default="fp64,fp32,split_bf16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--workload", default="full-disk", choices=["full-disk", "uneven-batch", "configs"],
                    help="full-disk: the 4096^2 sigma=15 pair (default); uneven-batch: 64 pairs "
                         "of sizes 256..2048 and sigma 4..20 in one graph launch; configs: the "
                         "128^2, 512^2 and 1024^2 cases with their rates and errors")
    ap.add_argument("--pairs", type=int, default=64)
    ap.add_argument("--no-autotune", action="store_true",
                    help="uneven-batch: time the untuned path rule only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; multi-GPU runs are not supported "
              "(across GPUs the tracker would take row strips of the image, one per GPU)", file=sys.stderr)
        sys.exit(2)

    from pyflct_amd import api, synth
    from pyflct_amd._lib import FlctHipError

    if args.workload == "uneven-batch":
        return uneven_batch(args)
    if args.workload == "configs":
        return configs_bench(args)

    log(f"bench: generating the {args.size}^2 pair")
    f1, f2, vtrue = synth.cfg4(n=args.size)
    nx, ny = f1.shape
    ctx = api.Context(0)
    nbytes = f1.nbytes
    d = {k: ctx.malloc(nbytes) for k in ("f1", "f2", "vx", "vy", "vm")}
    ctx.upload(d["f1"], f1)
    ctx.upload(d["f2"], f2)

    forms = [f for f in args.forms.split(",") if f]
    results = {}
    ref = None
    for form in forms:
        opts = ctx.make_opts(1.0, 1.0, args.sigma, precision=form)
        try:
            for _ in range(args.warmup):
                ctx.track_device(d["f1"], d["f2"], nx, ny, d["vx"], d["vy"], d["vm"], opts)
        except FlctHipError as e:
            log(f"bench: form {form} not available: {e}")
            continue
        log(f"bench: timing {form}, {args.steps} steps")
        kms, mms, rms_ = [], [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.track_device(d["f1"], d["f2"], nx, ny, d["vx"], d["vy"], d["vm"], opts)
            st = ctx.last_stats()
            kms.append(st["kernel_ms"])
            mms.append(st["main_ms"])
            rms_.append(st["refine_ms"])
        t1 = time.perf_counter()
        st = ctx.last_stats()
        ms = (t1 - t0) * 1e3 / args.steps
        vx = ctx.download(np.empty_like(f1), d["vx"])
        vy = ctx.download(np.empty_like(f1), d["vy"])
        results[form] = {"ms_per_step": ms, "tracked": st["tracked"],
                         "px_per_s": st["tracked"] / (ms * 1e-3),
                         "kernel_ms": float(np.mean(kms)), "main_ms": float(np.mean(mms)),
                         "refine_ms": float(np.mean(rms_)), "refined": st["refined"],
                         "path": ctx.last_path(), "vx": vx, "vy": vy}
        if form == "fp64":
            ref = (vx, vy)
        log(f"bench: {form}: {ms:.1f} ms/step, kernel {np.mean(kms):.1f} ms, "
            f"{results[form]['px_per_s'] / 1e6:.2f} Mpx/s")

    if "fp64" not in results:
        raise SystemExit("bench: the FP64 plain form failed")
    nbox = ctx.box_size(nx, ny, args.sigma)
    # LDS bandwidth of each form's tracking step: bytes counted by rocprofv3
    # (SQ_INSTS_LDS_LOAD_BANDWIDTH + SQ_INSTS_LDS_STORE_BANDWIDTH, 64 B units)
    # on this workload when the counter file holds them, beside the access
    # pattern model of the FFT kernels.  The counter file describes the
    # 4096^2, sigma = 15 pair only.
    pmc = {}
    if os.path.exists(PMC_FILE) and nbox == 64 and (nx, ny) == (4096, 4096):
        try:
            pmc = json.load(open(PMC_FILE))
        except (OSError, ValueError):
            pmc = {}
    lds = {}
    for form, r in results.items():
        counted = pmc.get(form, {}).get("lds_bytes_per_launch")
        model = None
        if r["path"] in ("fp64-fft", "fp32-fft"):  # the quad-FFT tracker's access pattern
            model = lds_bytes_fft(nbox, 16 if form == "fp64" else 8,
                                  2 if (form == "fp32" and 44 <= nbox <= 96) else 1) * r["tracked"]
        b = counted if counted else model
        if not b:
            continue
        tbs = b / (r["main_ms"] * 1e-3) / 1e12
        lds[form] = {"bytes_per_px": round(b / r["tracked"]), "source": "rocprofv3 counters" if counted
                     else "access-pattern model", "achieved_TBps": round(tbs, 2),
                     "peak_TBps": round(PEAK_LDS_TBS, 1), "frac": round(tbs / PEAK_LDS_TBS, 4)}
        if model and counted:
            lds[form]["model_bytes_per_px"] = round(model / r["tracked"])
    for form, r in results.items():
        r["rms_vel_err_vs_fp64"] = float(np.sqrt(np.mean((r["vx"] - ref[0]) ** 2 +
                                                         (r["vy"] - ref[1]) ** 2)))
        r["max_vel_err_vs_fp64"] = float(max(np.max(np.abs(r["vx"] - ref[0])),
                                             np.max(np.abs(r["vy"] - ref[1]))))

    # sampled parity of the FP64 form against the CPU oracle
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    rng = np.random.default_rng(1)
    pix = rng.choice(nx * ny, 64, replace=False)
    ox, oy, _ = oracle_lib.flct_pixels(f1, f2, 1.0, 1.0, args.sigma, pix)
    parity = float(max(np.max(np.abs(ref[0].ravel()[pix] - ox)),
                       np.max(np.abs(ref[1].ravel()[pix] - oy))))

    # headline: the split-BF16 form when it meets the 1e-4 px RMS bound
    head = "fp64"
    if "split_bf16" in results and results["split_bf16"]["rms_vel_err_vs_fp64"] <= 1e-4:
        head = "split_bf16"
    hr = results[head]

    # end to end through the host-buffer entry point: H2D of both images
    # (pageable float64), the tracking step, D2H of vx, vy, vm
    e2e = []
    for _ in range(2):
        t = time.perf_counter()
        ctx.track(f1, f2, 1.0, 1.0, args.sigma, precision=head)
        e2e.append((time.perf_counter() - t) * 1e3)
    end_to_end = {"form": head, "ms": round(e2e[-1], 3),
                  "px_per_s": round(hr["tracked"] / (e2e[-1] * 1e-3), 1),
                  "buffers": "host float64 (pageable): 2 images in, vx, vy, vm out"}
    log(f"bench: end to end {e2e[-1]:.1f} ms")
    dtype = {"fp64": "f64", "fp32": "f32", "split_bf16": "bf16x3+f32acc"}[head]

    # the kernel family of the path the headline form ran
    fam = {"fp64-rowseg": "k_track_rs", "fp32-rowseg": "k_track_rs"}.get(hr["path"], "k_track")
    if head == "fp64":
        peak, per_px, kname = PEAK_FP64_TFLOPS, w_fft(nbox), f"{fam}<{nbox // 4},double>"
    elif head == "fp32":
        peak, per_px, kname = PEAK_FP32_TFLOPS, w_fft(nbox), f"{fam}<{nbox // 4},float>"
    else:
        peak, per_px = PEAK_BF16_TFLOPS, w_gemm(nbox)
        kname = ("k_track_seg (interior columns and the two transposed edge strips)"
                 " + k_track_r2<2> (corners)"
                 if 36 <= nbox <= 64 and hr["path"] == "bf16-segment" else "k_track_mfma")
    seg_basis = head == "split_bf16" and 36 <= nbox <= 64
    if seg_basis:
        # the contractions the row-segment algorithm needs (radix-2 column
        # stage, box-row stage shared by 64 pixels, register-chained
        # inverse): a third of the dense-DFT count w_gemm, which stays in the
        # line as dense_dft_* for comparison
        nedge = nx * (nbox - 1)  # clamped edge columns, per-pixel kernel
        flops = seg_mfma_flops(nbox) * (hr["tracked"] - nedge) + r2_mfma_flops(nbox) * nedge
    else:
        flops = per_px * hr["tracked"]
    achieved = flops / (hr["main_ms"] * 1e-3) / 1e12
    traffic = pmc.get(head, {}).get("hbm_bytes_per_launch")
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": kname,
            "flops_per_px": round(flops / hr["tracked"]), "kernel_ms": round(hr["main_ms"], 3)}
    if seg_basis:
        dense = per_px * hr["tracked"] / (hr["main_ms"] * 1e-3) / 1e12
        roof["basis"] = "contractions of the row-segment algorithm (= executed MFMA flops)"
        roof["dense_dft_flops_per_px"] = per_px
        roof["dense_dft_frac"] = round(dense / peak, 4)

    # warp_frac2d of image 1 by the FP64 velocities (the NUFFT evaluation)
    ctx.upload(d["vx"], ref[0])
    ctx.upload(d["vy"], ref[1])
    wms = []
    for _ in range(1 + max(1, args.steps)):
        ctx.warp_device(d["f1"], d["vx"], d["vy"], d["vm"], nx, ny)
        wms.append(ctx.last_stats()["total_ms"])
    warp = {"workload": f"warp_frac2d of the {nx}x{ny} image 1 by the FP64 velocities",
            "method": ctx.last_warp_method(), "ms": round(float(np.mean(wms[1:])), 3),
            "px_per_s": round(nx * ny / (np.mean(wms[1:]) * 1e-3), 1)}
    log(f"bench: warp {nx}x{ny}: {warp['ms']:.2f} ms ({warp['method']})")

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle baseline")
        v, sample = cpu_baseline(f1, f2, args.sigma, args.cpu_budget)
        hc = host_cpu()
        cpu = {"value": round(v, 1), "unit": "tracked px/s", "cores": 1, "kind": "port",
               "sample": sample, "cpu_model": hc["cpu_model"], "nproc": hc["nproc"],
               "threads_available": hc["threads_available"],
               "fft": ("the port's own mixed-radix FFT; host libfftw3: "
                       + (hc["host_libfftw3"] or "none installed (ldconfig -p)"))}

    out = {
        "metric": METRIC,
        "value": round(hr["px_per_s"], 1),
        "unit": "tracked px/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(hr["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (numpy PCG64 seed 20261206 full-disk magnetogram pair)",
        "config": {"workload": f"full-disk pair {nx}x{ny}, sigma={args.sigma:g} "
                               f"({nbox}x{nbox} boxes), all pixels tracked",
                   "form": head, "box": nbox, "tracked_px": hr["tracked"]},
        "forms": {f: {"path": r["path"], "px_per_s": round(r["px_per_s"], 1),
                      "ms_per_step": round(r["ms_per_step"], 3),
                      "kernel_ms": round(r["kernel_ms"], 3), "main_ms": round(r["main_ms"], 3),
                      "refine_ms": round(r["refine_ms"], 3), "refined_px": r["refined"],
                      "rms_vel_err_vs_fp64": r["rms_vel_err_vs_fp64"],
                      "max_vel_err_vs_fp64": r["max_vel_err_vs_fp64"],
                      **({"roofline_fft": form_roofline(f, r, nbox)} if f in ("fp64", "fp32") else {})}
                  for f, r in results.items()},
        "fp64_vs_oracle_max_err_px": parity,
        "end_to_end": end_to_end,
        "lds": lds,
        "roofline": roof,
        "cpu_baseline": cpu,
        "warp": warp,
    }
    for k in d.values():
        ctx.free(k)
    print(json.dumps(out))


def uneven_batch(args):
    """64 image pairs, sizes uniform in {256, 512, 1024, 2048}^2, sigma
    uniform in 4-20, each a magnetogram warped by its own smooth random
    velocity field (8 modes, wavelengths 64-512 px, max |v| 1.5 px).  All
    pairs are tracked by one flct_hip_track_batch call (one graph launch) in
    FP64, FP32 and split-BF16; FP32 and split-BF16 are compared with FP64 and
    image 1 is warped by the split-BF16 velocities to report the residual
    against image 2."""
    from pyflct_amd import api, synth
    rng = np.random.default_rng(20261207)
    ctx = api.Context(0)
    pairs = []
    total_px = 0
    log(f"bench: generating {args.pairs} pairs")
    for k in range(args.pairs):
        n = int(rng.choice([256, 512, 1024, 2048]))
        sigma = float(rng.uniform(4.0, 20.0))
        f1, dx, dy = synth.uneven_inputs(n, seed=int(rng.integers(1 << 30)))
        d = {key: ctx.malloc(f1.nbytes) for key in ("f1", "f2", "dx", "dy")}
        ctx.upload(d["f1"], f1)
        ctx.upload(d["dx"], dx)
        ctx.upload(d["dy"], dy)
        ctx.warp_device(d["f1"], d["dx"], d["dy"], d["f2"], n, n)
        for form in ("fp64", "fp32", "split_bf16"):
            for o in ("vx", "vy", "vm"):
                d[f"{form}_{o}"] = ctx.malloc(f1.nbytes)
        pairs.append((n, sigma, d))
        total_px += n * n
    log(f"bench: {total_px / 1e6:.1f} Mpx in {len(pairs)} pairs; "
        f"sizes {sorted(collections.Counter(p[0] for p in pairs).items())}")
    res = {}
    for form in ("fp64", "fp32", "split_bf16"):
        desc = [(d["f1"], d["f2"], d[f"{form}_vx"], d[f"{form}_vy"], d[f"{form}_vm"], n, n,
                 ctx.make_opts(1.0, 1.0, sigma, precision=form)) for n, sigma, d in pairs]
        for _ in range(max(1, args.warmup)):
            ctx.track_batch_device(desc)  # the first call captures the graph
        ms = []
        for _ in range(args.steps):
            st = ctx.track_batch_device(desc)
            ms.append(ctx.last_stats()["kernel_ms"])
        tracked = sum(x["tracked"] for x in st)
        res[form] = {"ms": float(np.mean(ms)), "px_per_s": tracked / (np.mean(ms) * 1e-3),
                     "tracked": tracked, "refined": sum(x["refined"] for x in st)}
        log(f"bench: {form}: {np.mean(ms):.1f} ms per batch, {res[form]['px_per_s'] / 1e6:.1f} Mpx/s")
    # the autotuner: each (form, box size) is tuned on the largest pair of
    # that box size, then the batch is timed again with the tuned paths (a
    # tuning changes the context's paths, so the batch graph is captured anew)
    tuned = {}
    if not args.no_autotune:
        by_box = {}
        for n, sigma, d in pairs:
            nb = api.Context.box_size(n, n, sigma)
            if nb not in by_box or n > by_box[nb][0]:
                by_box[nb] = (n, sigma, d)
        for form in ("fp32", "split_bf16"):
            t0 = time.perf_counter()
            paths = {}
            for nb, (n, sigma, d) in sorted(by_box.items()):
                ctx.autotune(d["f1"], d["f2"], n, n, ctx.make_opts(1.0, 1.0, sigma, precision=form))
                sl, sp = ctx.last_shape()
                paths[nb] = ctx.last_path() + (f" slots {sl}" if sl else "") + (f" segment {sp}" if sp else "")
                log(f"bench: {form} box {nb} ({n}^2): {paths[nb]}")
            tune_s = time.perf_counter() - t0
            desc = [(d["f1"], d["f2"], d[f"{form}_vx"], d[f"{form}_vy"], d[f"{form}_vm"], n, n,
                     ctx.make_opts(1.0, 1.0, sigma, precision=form)) for n, sigma, d in pairs]
            for _ in range(max(1, args.warmup)):
                ctx.track_batch_device(desc)
            ms = []
            for _ in range(args.steps):
                st = ctx.track_batch_device(desc)
                ms.append(ctx.last_stats()["kernel_ms"])
            tracked = sum(x["tracked"] for x in st)
            tuned[form] = {"ms": float(np.mean(ms)), "px_per_s": tracked / (np.mean(ms) * 1e-3),
                           "refined": sum(x["refined"] for x in st), "tune_s": tune_s,
                           "paths": {str(k): v for k, v in paths.items()}}
            log(f"bench: {form} tuned ({tune_s:.1f} s of tuning): {np.mean(ms):.1f} ms per batch, "
                f"{tuned[form]['px_per_s'] / 1e6:.1f} Mpx/s; paths {paths}")
    # with the autotuner on, the FP32 and split-BF16 velocities compared
    # below are the tuned batches'
    # the warp of every image 1 by its split-BF16 velocities, timed (device
    # time per call, summed over the pairs), then the same for one 2048^2
    # pair with the reference's direct per-pixel sum
    wms = 0.0
    for n, sigma, d in pairs:
        ctx.warp_device(d["f1"], d["split_bf16_vx"], d["split_bf16_vy"], d["dx"], n, n)
        wms += ctx.last_stats()["total_ms"]
    wmethod = ctx.last_warp_method()
    direct = None
    big = [p_ for p_ in pairs if p_[0] == 2048]
    if big:
        n, sigma, d = big[0]
        ctx.set_warp_method("direct")
        try:
            ctx.warp_device(d["f1"], d["split_bf16_vx"], d["split_bf16_vy"], d["dy"], n, n)
            direct = round(ctx.last_stats()["total_ms"], 3)
        finally:
            ctx.set_warp_method("auto")
    log(f"bench: warps of all pairs {wms:.1f} ms ({wmethod}); one 2048^2 direct sum {direct} ms")
    # errors against FP64 and the warp residual of the split-BF16 velocities
    sq = {"fp32": 0.0, "split_bf16": 0.0}
    mx = {"fp32": 0.0, "split_bf16": 0.0}
    rres, rbase = [], []
    for n, sigma, d in pairs:
        ref = [ctx.download(np.empty((n, n)), d[f"fp64_{o}"]) for o in ("vx", "vy")]
        for form in sq:
            g = [ctx.download(np.empty((n, n)), d[f"{form}_{o}"]) for o in ("vx", "vy")]
            e = np.hypot(g[0] - ref[0], g[1] - ref[1])
            sq[form] += float(np.sum(e ** 2))
            mx[form] = max(mx[form], float(e.max()))
        ctx.warp_device(d["f1"], d["split_bf16_vx"], d["split_bf16_vy"], d["dx"], n, n)
        w = ctx.download(np.empty((n, n)), d["dx"])
        a = ctx.download(np.empty((n, n)), d["f1"])
        b = ctx.download(np.empty((n, n)), d["f2"])
        m = slice(n // 8, n - n // 8)
        rres.append(np.sqrt(np.mean((w[m, m] - b[m, m]) ** 2)))
        rbase.append(np.sqrt(np.mean((a[m, m] - b[m, m]) ** 2)))
    for n, sigma, d in pairs:
        for p_ in d.values():
            ctx.free(p_)
    hr = tuned.get("split_bf16", res["split_bf16"])
    out = {
        "metric": "uneven batch tracked px/s (64 pairs, 256..2048, sigma 4..20), split-BF16",
        "value": round(hr["px_per_s"], 1),
        "unit": "tracked px/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(hr["ms"], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3+f32acc",
        "data": "synthetic (PCG64 seed 20261207: magnetograms warped by random smooth fields)",
        "config": {"workload": "uneven batch of 64 pairs in one graph launch", "pairs": len(pairs),
                   "total_px": total_px},
        "forms": {f: {"px_per_s": round(r["px_per_s"], 1), "ms_per_batch": round(r["ms"], 3),
                      "refined_px": r["refined"]} for f, r in res.items()},
        "autotuned": {f: {"px_per_s": round(r["px_per_s"], 1), "ms_per_batch": round(r["ms"], 3),
                          "refined_px": r["refined"], "tuning_s": round(r["tune_s"], 2),
                          "path_by_box": r["paths"]} for f, r in tuned.items()},
        "rms_vel_err_vs_fp64": {f: float(np.sqrt(sq[f] / total_px)) for f in sq},
        "max_vel_err_vs_fp64": mx,
        "warp_residual_rms": {"split_bf16_velocities": float(np.sqrt(np.mean(np.square(rres)))),
                              "unwarped": float(np.sqrt(np.mean(np.square(rbase))))},
        "warp": {"method": wmethod, "ms_all_pairs": round(wms, 3),
                 "px_per_s": round(total_px / (wms * 1e-3), 1),
                 "direct_sum_ms_one_2048_pair": direct},
        "tracking_plus_warp_ms": round(hr["ms"] + wms, 3),
    }
    print(json.dumps(out))


def configs_bench(args):
    """The three smaller cases, each timed on resident buffers with its error
    stated: 128^2 blobs shifted by (0.3, -0.2), "1 1 8" (FP64); 512^2
    magnetogram under a 0.01 degree rotation, "1 1 6 -t 0.1 -k 0.5 -b"
    (FP64); 1024^2 float32 magnetogram under a smooth random field,
    "1 1 10 -s 4i", in FP64 and in the FP32 form through one graph launch
    (flct_hip_track_batch).  FP64 is compared with the CPU oracle (every
    pixel at 128^2, 2000 sampled tracked pixels otherwise), FP32 with FP64."""
    from pyflct_amd import api, synth
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    ctx = api.Context(0)
    rng = np.random.default_rng(5)
    steps = max(1, args.steps)

    def resident(f1, f2):
        d = {k: ctx.malloc(f1.nbytes) for k in ("f1", "f2", "vx", "vy", "vm")}
        ctx.upload(d["f1"], f1)
        ctx.upload(d["f2"], f2)
        return d

    def timed(d, nx, ny, opts):
        ms = []
        for _ in range(1 + steps):
            ctx.track_device(d["f1"], d["f2"], nx, ny, d["vx"], d["vy"], d["vm"], opts)
            ms.append(ctx.last_stats()["kernel_ms"])
        tracked = ctx.last_stats()["tracked"]
        out = [ctx.download(np.empty((nx, ny)), d[k]) for k in ("vx", "vy", "vm")]
        return float(np.mean(ms[1:])), tracked, out

    def vs_oracle(f1, f2, sigma, v, pix, **kw):
        ox, oy, om = oracle_lib.flct_pixels(f1, f2, 1.0, 1.0, sigma, pix, **kw)
        return float(max(np.max(np.abs(v[0].ravel()[pix] - ox)), np.max(np.abs(v[1].ravel()[pix] - oy))))

    res = {}
    # 128^2, "1 1 8"
    f1, f2 = synth.cfg1()
    d = resident(f1, f2)
    ms, tr, v = timed(d, 128, 128, ctx.make_opts(1.0, 1.0, 8.0))
    o = oracle_lib.flct(f1, f2, 1.0, 1.0, 8.0)
    res["blobs_128_sigma8"] = {"form": "fp64", "kernel_ms": round(ms, 4), "tracked_px": tr,
                               "tracked_px_per_s": round(tr / (ms * 1e-3), 1),
                               "max_err_vs_oracle_px": float(max(np.max(np.abs(v[0] - o[0])),
                                                                 np.max(np.abs(v[1] - o[1])))),
                               "vm_equal": bool(np.array_equal(v[2], o[2])),
                               "mean_v_interior": [float(np.mean(v[0][32:96, 32:96])),
                                                   float(np.mean(v[1][32:96, 32:96]))]}
    for k in d.values():
        ctx.free(k)
    log(f"bench: 128^2 sigma 8: {ms:.3f} ms")
    # 512^2, "1 1 6 -t 0.1 -k 0.5 -b"
    f1, dx, dy = synth.cfg2_inputs()
    f2 = ctx.warp(f1, dx, dy)
    d = resident(f1, f2)
    kw = dict(thresh=0.1, kr=0.5, bias=True)
    ms, tr, v = timed(d, 512, 512, ctx.make_opts(1.0, 1.0, 6.0, **kw))
    tracked_idx = np.flatnonzero(v[2].ravel() == 1.0)
    pix = rng.choice(tracked_idx, min(2000, tracked_idx.size), replace=False)
    res["magnetogram_512_sigma6_t0.1_k0.5_b"] = {
        "form": "fp64", "kernel_ms": round(ms, 4), "tracked_px": tr,
        "tracked_px_per_s": round(tr / (ms * 1e-3), 1),
        "max_err_vs_oracle_px": vs_oracle(f1, f2, 6.0, v, pix, **kw), "oracle_sample_px": int(pix.size)}
    for k in d.values():
        ctx.free(k)
    log(f"bench: 512^2 sigma 6 -t 0.1 -k 0.5 -b: {ms:.3f} ms, {tr} tracked")
    # 1024^2 float32, "1 1 10 -s 4i": FP64, then FP32 through one graph launch
    f1, dx, dy = synth.cfg3_inputs()
    f1 = f1.astype(np.float32).astype(np.float64)
    f2 = ctx.warp(f1, dx, dy).astype(np.float32).astype(np.float64)
    d = resident(f1, f2)
    kw = dict(skip=4, interp=True)
    ms64, tr, v64 = timed(d, 1024, 1024, ctx.make_opts(1.0, 1.0, 10.0, **kw))
    lat = np.flatnonzero(v64[2].ravel() == 1.0)
    pix = rng.choice(lat, min(2000, lat.size), replace=False)
    err64 = vs_oracle(f1, f2, 10.0, v64, pix, skip=4)
    o32 = {k: ctx.malloc(f1.nbytes) for k in ("vx", "vy", "vm")}
    desc = [(d["f1"], d["f2"], o32["vx"], o32["vy"], o32["vm"], 1024, 1024,
             ctx.make_opts(1.0, 1.0, 10.0, precision="fp32", **kw))]
    ms32 = []
    for _ in range(1 + steps):
        st = ctx.track_batch_device(desc)
        ms32.append(ctx.last_stats()["kernel_ms"])
    v32 = [ctx.download(np.empty((1024, 1024)), o32[k]) for k in ("vx", "vy", "vm")]
    ms32 = float(np.mean(ms32[1:]))
    e = np.hypot(v32[0] - v64[0], v32[1] - v64[1])
    m1, mh = v64[2] == 1.0, v64[2] == 0.5
    res["magnetogram_1024_f32_sigma10_s4i"] = {
        "fp64": {"kernel_ms": round(ms64, 4), "tracked_px": tr,
                 "tracked_px_per_s": round(tr / (ms64 * 1e-3), 1),
                 "output_px_per_s": round(1024 * 1024 / (ms64 * 1e-3), 1),
                 "max_err_vs_oracle_px": err64, "oracle_sample_px": int(pix.size)},
        "fp32_graph": {"kernel_ms": round(ms32, 4), "tracked_px": st[0]["tracked"],
                       "refined_px": st[0]["refined"],
                       "tracked_px_per_s": round(st[0]["tracked"] / (ms32 * 1e-3), 1),
                       "output_px_per_s": round(1024 * 1024 / (ms32 * 1e-3), 1),
                       "vm_equal_fp64": bool(np.array_equal(v32[2], v64[2])),
                       "rms_err_vs_fp64_tracked": float(np.sqrt(np.mean(e[m1] ** 2))),
                       "max_err_vs_fp64_tracked": float(e[m1].max()),
                       "rms_err_vs_fp64_filled": float(np.sqrt(np.mean(e[mh] ** 2))),
                       "max_err_vs_fp64_filled": float(e[mh].max())}}
    for k in list(d.values()) + list(o32.values()):
        ctx.free(k)
    log(f"bench: 1024^2 sigma 10 -s 4i: FP64 {ms64:.3f} ms, FP32 graph {ms32:.3f} ms")
    print(json.dumps({"metric": "tracked px/s and errors of the 128^2, 512^2 and 1024^2 cases",
                      "n_gpus": 1, "steps": steps, "data": "synthetic (PCG64, see pyflct_amd/synth.py)",
                      "configs": res}))


if __name__ == "__main__":
    main()
