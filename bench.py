# SPDX-License-Identifier: Apache-2.0
"""Benchmark of the decomposition hot path on one MI355X: eda.decompose of 64 float32 images of
512 x 512 (rings, warped stripes, two Gaussian blobs and noise; per-image seeds), the input already
resident in HBM as a PyTorch-ROCm tensor, default parameters.

A step is one eda.decompose call over the whole batch (every mode of every image, rows collected
into a device tensor). W untimed warm-up steps, then K timed steps bracketed by a device
synchronization. Prints one JSON line on stdout (progress goes to stderr):

  value      Msamples/s = 64 * 512 * 512 / (seconds per step)
  roofline   the dominant kernel, the cube filter's vHGW pass (one pass = one launch of k_vhgw_rows
             or k_vhgw, or the one or two launches of k_box of a box-sum pass): algorithmic bytes
             per pass (its input field(s) read once, output field(s) written once, over the running
             elements) / its average duration, both measured live with HIP events on the library's
             stream over the timed steps, against 8 TB/s; `traffic` the HBM bytes per pass from
             rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/), null when not measured
  cpu_baseline  the oracle (a C++ restatement of the method, FP64, one image per host thread)
             on a bounded sample of the same workload, on this host's cores

With --monogenic the same batch goes through eda.monogenic instead, and the roofline is the DCT
axis pass (the fast transform of the 512-sample axes: one read and one write of the field).

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--monogenic]
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (ROOT / "eda-cuda_amd", ROOT / "oracle", ROOT / "tests"):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))

METRIC = "Msamples/s decomposed (64×512² FP32), fwd and fwd+bwd, % of HBM roofline"
HBM_PEAK_GBS = 8000.0
B, N = 64, 512


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """HBM bytes per vHGW pass measured by rocprofv3 (profiles/vhgw_pass_64x512x512_f32.json), or
    None."""
    p = ROOT / "profiles" / "vhgw_pass_64x512x512_f32.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("traffic_bytes_per_pass")
    except Exception:
        return None


COPY_PROBE_BYTES = 1 << 30  # the known-size copy of --copy-probe: 1 GiB read + 1 GiB written


def _counter_rows(directory):
    """(kernel name, counter name, value) of every counter_collection.csv under directory."""
    import csv

    for path in sorted(Path(directory).rglob("*counter_collection.csv")):
        with open(path, newline="") as f:
            for row in csv.DictReader(f):
                yield row.get("Kernel_Name", ""), row.get("Counter_Name", ""), float(row.get("Counter_Value", 0.0))


def pmc_summary(fetch_dir, write_dir, out_json):
    """HBM traffic per vHGW pass from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE, KiB
    per dispatch) of `bench.py --steps 1 --warmup 0 --no-cpu --copy-probe`, each scaled by what the
    counter reads for the probe's copy kernel (torch.neg over 1 GiB: 1 GiB read, 1 GiB written).
    The pass kernels' KiB are summed and divided by the number of passes: every k_vhgw_rows or
    k_vhgw dispatch is a pass, and a box-sum pass is one k_box<..., false> dispatch and at most one
    k_box<..., true>."""
    def averages(directory, counter):
        total, passes, cp = 0.0, 0, []
        for name, cname, value in _counter_rows(directory):
            if cname != counter:
                continue
            if "k_vhgw" in name or "k_box" in name:
                total += value
                if "k_vhgw" in name or "false>" in name:
                    passes += 1
            elif "neg" in name:
                cp.append(value)
        if not passes or not cp:
            raise SystemExit(f"bench: no {counter} rows for the vHGW pass and the copy probe under {directory}")
        return total / passes, max(cp), passes

    f_vh, f_cp, n_f = averages(fetch_dir, "FETCH_SIZE")
    w_vh, w_cp, n_w = averages(write_dir, "WRITE_SIZE")
    f_scale = COPY_PROBE_BYTES / (f_cp * 1024.0)
    w_scale = COPY_PROBE_BYTES / (w_cp * 1024.0)
    traffic = f_vh * 1024.0 * f_scale + w_vh * 1024.0 * w_scale
    res = {
        "kernel": "vHGW pass (k_vhgw_rows, k_vhgw, k_box)",
        "workload": "eda.decompose fwd, 64 x 512 x 512 float32 (one call)",
        "passes": n_f,
        "fetch_kib_per_pass": round(f_vh, 1),
        "write_kib_per_pass": round(w_vh, 1),
        "copy_probe_fetch_kib": round(f_cp, 1),
        "copy_probe_write_kib": round(w_cp, 1),
        "fetch_scale": round(f_scale, 4),
        "write_scale": round(w_scale, 4),
        "traffic_bytes_per_pass": round(traffic),
    }
    Path(out_json).write_text(json.dumps(res, indent=1) + "\n")
    print(json.dumps(res))
    return 0


def copy_probe(torch, log_fn):
    """One elementwise copy kernel of known size (torch.neg over 1 GiB of float32) for scaling the
    HBM byte counters."""
    n = COPY_PROBE_BYTES // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    torch.neg(a, out=b)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    log_fn(f"bench: copy probe: {2 * COPY_PROBE_BYTES / dt / 1e9:.0f} GB/s (torch.neg over 1 GiB, wall clock)")
    del a, b


def cpu_baseline():
    import numpy as np

    import oracle
    import workloads

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    images = min(B, 2 * threads)
    x = workloads.rings_stripes_blobs(images, N).astype(np.float64)
    log(f"bench: cpu baseline: the FP64 oracle on {images} image(s) of {N}x{N}, {threads} thread(s)")
    t0 = time.perf_counter()
    oracle.decompose(x, dim=2, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": round(images * N * N / dt / 1e6, 6),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"images 0..{images - 1} of the 64 ({N}x{N}, float64 oracle, OpenMP over images), {dt:.1f} s",
    }


MONO_METRIC = "Msamples/s monogenic signal (64×512² FP32)"


def bench_monogenic(args, x, lib, torch) -> dict:
    """eda.monogenic of the same 64 x 512 x 512 float32 batch (input resident in HBM): throughput,
    the DCT axis passes' live rate (one read and one write of the field per pass, timed with HIP
    events on the library's stream) against the HBM peak, and the FP64 oracle's rate on a bounded
    sample."""
    import numpy as np

    import eda
    from eda import _native

    for _ in range(args.warmup):
        r = eda.monogenic(x, axes=(1, 2))
        torch.cuda.synchronize()
        del r
    err = _native.errbuf()
    lib.eda_timer_enable(2)
    _native.check(lib.eda_timer_read(ctypes.byref(ctypes.c_double()), ctypes.byref(ctypes.c_int64()),
                                     ctypes.byref(ctypes.c_double()), err, len(err)), err)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = eda.monogenic(x, axes=(1, 2))
        del r
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    ms, launches, nbytes = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    _native.check(lib.eda_timer_read(ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(nbytes), err, len(err)),
                  err)
    lib.eda_timer_enable(0)
    avg_s = ms.value * 1e-3 / max(launches.value, 1)
    per_launch = nbytes.value / max(launches.value, 1)
    achieved = per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    log(f"bench: monogenic: {args.steps} steps, {dt * 1e3:.1f} ms per step; DCT axis passes: {launches.value}, "
        f"{ms.value:.1f} ms in total")
    out = {
        "metric": MONO_METRIC,
        "value": round(B * N * N / dt / 1e6, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "eda.monogenic, 64 x 512 x 512 float32 images (the decomposition benchmark's "
                               "inputs), input resident in HBM", "batch": B, "shape": [N, N]},
        "roofline": {
            "bound": "hbm",
            "kernel": "DCT axis pass (k_dct_fft: one axis of one transform per launch)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": round(per_launch),
            "avg_launch_us": round(avg_s * 1e6, 2),
            "launches_per_step": launches.value // max(args.steps, 1),
            "share_of_step": round(ms.value * 1e-3 / max(args.steps, 1) / dt, 4),
        },
    }
    if not args.no_cpu:
        import oracle
        import workloads
        from concurrent.futures import ThreadPoolExecutor

        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count() or 1
        threads = max(1, min(16, cores))
        images = threads
        xs = workloads.rings_stripes_blobs(images, N).astype(np.float64)
        log(f"bench: monogenic cpu baseline: the FP64 oracle on {images} image(s), {threads} thread(s)")
        t = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            list(pool.map(oracle.monogenic, list(xs)))
        cdt = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": round(images * N * N / cdt / 1e6, 6),
            "unit": "Msamples/s",
            "cores": threads,
            "kind": "port",
            "sample": f"images 0..{images - 1} of the 64 ({N}x{N}, float64 oracle, one image per thread), {cdt:.1f} s",
        }
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one copy kernel of known size (for scaling rocprofv3 byte counters)")
    ap.add_argument("--monogenic", action="store_true",
                    help="measure eda.monogenic on the same batch instead of eda.decompose")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"),
                    help="summarize two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary)
    if args.gpus != 1:
        print(
            "bench.py: measures one GPU only (the library runs on one GPU per call; batch elements "
            "would shard across GPUs with no exchange); --gpus must be 1",
            file=sys.stderr,
        )
        return 2

    import numpy as np
    import torch

    import eda
    import workloads
    from eda import _native

    assert torch.cuda.is_available(), "bench: no GPU"
    lib = _native.core()
    log("bench: generating 64 x 512 x 512 float32 inputs")
    x = torch.from_numpy(workloads.rings_stripes_blobs(B, N, np.float32)).to("cuda:0")
    torch.cuda.synchronize()
    if args.monogenic:
        print(json.dumps(bench_monogenic(args, x, lib, torch)), flush=True)
        return 0

    info = None
    for i in range(args.warmup):
        t = time.perf_counter()
        y, info = eda.decompose(x, axes=(1, 2), return_info=True)
        torch.cuda.synchronize()
        log(f"bench: warm-up {i}: {time.perf_counter() - t:.3f} s, n_imfs {sorted(set(info.n_imfs.tolist()))}")
        del y

    err = _native.errbuf()
    lib.eda_timer_enable(1)
    _native.check(lib.eda_timer_read(ctypes.byref(ctypes.c_double()), ctypes.byref(ctypes.c_int64()),
                                     ctypes.byref(ctypes.c_double()), err, len(err)), err)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        y, info = eda.decompose(x, axes=(1, 2), return_info=True)
        del y
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    ms = ctypes.c_double()
    launches = ctypes.c_int64()
    nbytes = ctypes.c_double()
    _native.check(lib.eda_timer_read(ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(nbytes), err, len(err)),
                  err)
    lib.eda_timer_enable(0)
    log(f"bench: {args.steps} steps, {dt * 1e3:.1f} ms per step; vHGW pass: {launches.value} passes, "
        f"{ms.value:.1f} ms in total")

    if args.copy_probe:
        copy_probe(torch, log)

    # the host-input rate (numpy in, numpy out: PCIe transfers included); reported on stderr only
    xh = x.cpu().numpy()
    t = time.perf_counter()
    yh = eda.decompose(xh, axes=(1, 2))
    host_dt = time.perf_counter() - t
    log(f"bench: host input, PCIe-inclusive: {B * N * N / host_dt / 1e6:.1f} Msamples/s "
        f"({host_dt * 1e3:.1f} ms, output {yh.nbytes / 2**20:.0f} MiB)")
    del yh

    value = B * N * N / dt / 1e6
    per_launch_bytes = nbytes.value / max(launches.value, 1)
    avg_s = ms.value * 1e-3 / max(launches.value, 1)
    achieved = per_launch_bytes / avg_s / 1e9 if avg_s > 0 else 0.0
    roofline = {
        "bound": "hbm",
        "kernel": "vHGW pass (k_vhgw_rows, k_vhgw, k_box)",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(),
        "algorithmic_bytes_per_pass": round(per_launch_bytes),
        "avg_pass_us": round(avg_s * 1e6, 2),
        "passes_per_step": launches.value // max(args.steps, 1),
        "share_of_step": round(ms.value * 1e-3 / max(args.steps, 1) / dt, 4),
    }
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "eda.decompose fwd, 64 x 512 x 512 float32 images (rings + warped stripes + "
                        "2 Gaussian blobs + N(0, 0.05^2) noise, seed = image index), default parameters "
                        "(sifts 3), input resident in HBM",
            "batch": B,
            "shape": [N, N],
            "pass": "fwd",
            "n_imfs": sorted(set(info.n_imfs.tolist())) if info is not None else None,
        },
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
