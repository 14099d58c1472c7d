"""Sweep of the tracker's slots per block (pixels tracked together by one
block of k_track) over every box size.

    python -m pyflct_amd.tune.sweep_slots build [--slots 1,2,3,4,6,8]   # CPU: variant libraries
    python -m pyflct_amd.tune.sweep_slots run [--size 2048] > sweep.txt  # GPU: time each variant
    python -m pyflct_amd.tune.sweep_slots table sweep.txt                # the tuned_slots() table

Each variant is the whole library built with -DFLCT_TRACK_SLOTS=p (every
box size and precision at p slots, reduced to what fits a CU) into
pyflct_amd/tune/_variants/.  `run` times the default build and every
variant, one subprocess per library (FLCT_HIP_LIB), on the uneven-batch
generator's image pair for sigma = 4..23 (boxes 16..96), FP64 and FP32,
every pixel tracked: the device time of the tracking kernel, best of two.
`table` prints, per precision and box, the count that beat the default by
the given margin (the switch in track_kernels.hpp's tuned_slots()).
"""
import argparse
import collections
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
VAR = os.path.join(HERE, "_variants")


def build(slots, jobs):
    for p in slots:
        obj = os.path.join(VAR, f"obj_p{p}")
        lib = os.path.join(VAR, f"libflct_hip_p{p}.so")
        subprocess.run(["make", f"-j{jobs}", "-C", os.path.join(PKG, "csrc"), f"OBJ={obj}", f"LIB={lib}",
                        f"TUNE_FLAGS=-DFLCT_TRACK_SLOTS={p}"], check=True)
        print(lib, flush=True)


def time_one(size):
    """Runs in a child process with FLCT_HIP_LIB set: one line per case."""
    sys.path.insert(0, ROOT)
    from pyflct_amd import api, synth
    ctx = api.Context(0)
    f1, dx, dy = synth.uneven_inputs(size, seed=5)
    d = {k: ctx.malloc(f1.nbytes) for k in ("f1", "f2", "dx", "dy", "vx", "vy", "vm")}
    ctx.upload(d["f1"], f1)
    ctx.upload(d["dx"], dx)
    ctx.upload(d["dy"], dy)
    ctx.warp_device(d["f1"], d["dx"], d["dy"], d["f2"], size, size)
    name = os.path.basename(os.environ.get("FLCT_HIP_LIB", "default"))
    for sigma in range(4, 24):
        for form in ("fp64", "fp32"):
            o = ctx.make_opts(1.0, 1.0, sigma, precision=form)
            ctx.track_device(d["f1"], d["f2"], size, size, d["vx"], d["vy"], d["vm"], o)
            ts = []
            for _ in range(2):
                ctx.track_device(d["f1"], d["f2"], size, size, d["vx"], d["vy"], d["vm"], o)
                ts.append(ctx.last_stats()["main_ms"])
            print(f"{name} {form} {ctx.box_size(size, size, sigma)} {min(ts):.3f}", flush=True)


def run(size):
    libs = [None] + sorted(os.path.join(VAR, f) for f in os.listdir(VAR) if f.endswith(".so"))
    for lib in libs:
        env = dict(os.environ)
        env.pop("FLCT_HIP_LIB", None)
        if lib:
            env["FLCT_HIP_LIB"] = lib
        print(f"# {lib or 'default build'}", file=sys.stderr, flush=True)
        subprocess.run([sys.executable, "-u", __file__, "_time", "--size", str(size)], env=env,
                       check=True, timeout=600)


def table(path, margin):
    t = collections.defaultdict(dict)
    for line in open(path):
        f = line.split()
        if len(f) == 4 and not line.startswith("#"):
            t[(f[1], int(f[2]))][f[0]] = float(f[3])
    for (form, n), row in sorted(t.items()):
        if "default" not in row:
            continue
        ms, best = min((v, k) for k, v in row.items())
        if best != "default" and row["default"] >= (1. + margin) * ms:
            p = best.split("_p")[-1].split(".")[0]
            print(f"{form} box {n} (M = {n // 4}): {p} slots, {row['default']:.1f} -> {ms:.1f} ms")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["build", "run", "table", "_time"])
    ap.add_argument("path", nargs="?")
    ap.add_argument("--slots", default="1,2,3,4,6,8")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--margin", type=float, default=0.04)
    a = ap.parse_args()
    if a.cmd == "build":
        build([int(x) for x in a.slots.split(",")], a.jobs)
    elif a.cmd == "run":
        run(a.size)
    elif a.cmd == "_time":
        time_one(a.size)
    else:
        table(a.path, a.margin)


if __name__ == "__main__":
    main()
