# SPDX-License-Identifier: Apache-2.0
"""Benchmark of the decomposition hot path on one MI355X: eda.decompose of 64 float32 images of
512 x 512 (rings, warped stripes, two Gaussian blobs and noise; per-image seeds), the input already
resident in HBM as a PyTorch-ROCm tensor, default parameters.

A step is one eda.decompose call over the whole batch (every mode of every image, rows collected
into a device tensor). W untimed warm-up steps, then K timed steps bracketed by a device
synchronization, with no instrumentation (the library's kernel timer is off). Prints one JSON line on
stdout (progress goes to stderr):

  value      Msamples/s = 64 * 512 * 512 / (seconds per step over the K timed steps)
  roofline   the dominant kernel family, the cube filter's vHGW pass (k_vhgw_rows, k_vhgw_stream,
             k_vhgw, the two k_box launches and the fused 2-D cube k_cube2): algorithmic bytes per
             launch (its input field(s) read once, output field(s) written once, over the running
             elements) / its average duration, both
             measured live with HIP events on the library's stream over two extra, untimed steps
             (the ledger), against 8 TB/s; `kernels` the same figures per kernel (also the rung
             gather, the extrema tiles and the DCT passes); `traffic` the HBM bytes per launch from
             rocprofv3 FETCH_SIZE + WRITE_SIZE (profiles/), null when not measured
  per_item   the plain per-item form: the same 64 images one call per image, and the batched
             call's speed-up over it
  fwd_bwd    one step of the differentiation form: the forward with its tape kept, the loss
             gradient and the hand-written adjoint back to the input
  cpu_baseline  the oracle (a C++ restatement of the method, FP64, OpenMP over images) on the whole
             batch, on this host's cores (at most 16), the median of 3 runs when a run is short

With --monogenic the same batch goes through eda.monogenic instead, and the roofline is the DCT
axis pass (the fast transform of the 512-sample axes: one read and one write of the field).
--precision and --batching measure the precision workload (4 x 256^3 volumes, exact against
split-BF16 contractions) and the ragged batching workload (256 uneven 1-D signals and 32 images,
per field against eda.decompose_many) instead.

Usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1] [--no-cpu] [--no-bwd] [--no-per-item] [--fwd-bwd-only] [--monogenic]
       [--precision] [--batching]
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
    """HBM bytes per vHGW-pass launch measured by rocprofv3 (profiles/vhgw_pass_64x512x512_f32.json),
    or None."""
    p = ROOT / "profiles" / "vhgw_pass_64x512x512_f32.json"
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text()).get("traffic_bytes_per_launch")
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
    """HBM traffic per vHGW-pass launch from two rocprofv3 --pmc passes (FETCH_SIZE, then WRITE_SIZE,
    KiB per dispatch) of `bench.py --steps 1 --warmup 0 --no-cpu --copy-probe`, each scaled by what
    the counter reads for the probe's copy kernel (torch.neg over 1 GiB: 1 GiB read, 1 GiB written).
    The KiB of every k_vhgw_rows, k_vhgw_stream, k_vhgw, k_box and k_cube2 dispatch are summed and divided by their
    number."""
    def averages(directory, counter):
        total, launches, cp = 0.0, 0, []
        for name, cname, value in _counter_rows(directory):
            if cname != counter:
                continue
            if "k_vhgw" in name or "k_box" in name or "k_cube2" in name:
                total += value
                launches += 1
            elif "neg" in name:
                cp.append(value)
        if not launches or not cp:
            raise SystemExit(f"bench: no {counter} rows for the vHGW pass and the copy probe under {directory}")
        return total / launches, max(cp), launches

    f_vh, f_cp, n_f = averages(fetch_dir, "FETCH_SIZE")
    w_vh, w_cp, n_w = averages(write_dir, "WRITE_SIZE")
    f_scale = COPY_PROBE_BYTES / (f_cp * 1024.0)
    w_scale = COPY_PROBE_BYTES / (w_cp * 1024.0)
    traffic = f_vh * 1024.0 * f_scale + w_vh * 1024.0 * w_scale
    res = {
        "kernel": "vHGW pass (k_vhgw_rows, k_vhgw, k_vhgw_stream, k_box, k_cube2)",
        "workload": "eda.decompose fwd, 64 x 512 x 512 float32 (one call)",
        "launches": n_f,
        "fetch_kib_per_launch": round(f_vh, 1),
        "write_kib_per_launch": round(w_vh, 1),
        "copy_probe_fetch_kib": round(f_cp, 1),
        "copy_probe_write_kib": round(w_cp, 1),
        "fetch_scale": round(f_scale, 4),
        "write_scale": round(w_scale, 4),
        "traffic_bytes_per_launch": round(traffic),
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


def host_threads():
    """The host threads the CPU baseline uses and the cores this process may run on: every core of
    the affinity mask, at most 16 (one GPU's share of the host)."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    return max(1, min(16, cores)), cores


def cpu_baseline():
    """The FP64 oracle on the whole 64-image batch (OpenMP over images); the median of 3 runs when
    one run takes under 40 s, else one run."""
    import numpy as np

    import oracle
    import workloads

    threads, cores = host_threads()
    x = workloads.rings_stripes_blobs(B, N).astype(np.float64)
    log(f"bench: cpu baseline: the FP64 oracle on {B} images of {N}x{N}, {threads} thread(s) ({cores} cores in the "
        f"affinity mask)")
    times = []
    while len(times) < 3:
        t0 = time.perf_counter()
        oracle.decompose(x, dim=2, threads=threads)
        times.append(time.perf_counter() - t0)
        log(f"bench: cpu baseline run {len(times)}: {times[-1]:.1f} s")
        if times[0] >= 40.0:
            break
    dt = sorted(times)[len(times) // 2]
    return {
        "value": round(B * N * N / dt / 1e6, 6),
        "unit": "Msamples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"all {B} images ({N}x{N}, float64 oracle, OpenMP over images, {threads} threads of the "
                  f"{cores} cores in the affinity mask), median of {len(times)} run(s): "
                  + ", ".join(f"{t:.1f}" for t in times) + " s",
    }


PREC_METRIC = "Msamples/s decomposed (4×256³ FP32), exact vs split-BF16 DCT, error vs FP64 per mode"


def bench_precision(args, lib, torch) -> dict:
    """The precision workload: 4 x 256^3 float32 plane-wave volumes decomposed with the exact DCT /
    Riesz transforms and with split-BF16 (3-term) MFMA contractions; per path the call time (median
    of the timed steps), the DCT axis passes' live rate (HIP events; algorithmic bytes: one read and
    one write of the field; flops 2 n per sample and axis, three products each for split-BF16) and
    the per-mode relative L2 error against the float64 decomposition of the same volumes."""
    import numpy as np

    import eda
    import workloads
    from eda import _native

    x = torch.from_numpy(workloads.plane_wave_volumes(4, 256)).to("cuda:0")
    y64, i64 = eda.decompose(x.double(), axes=(1, 2, 3), return_info=True)
    n_samples = x.numel()
    err = _native.errbuf()
    buf = ctypes.create_string_buffer(4096)
    out = {}
    for dct in ("exact", "bf16x3"):
        for _ in range(args.warmup):
            eda.decompose(x, axes=(1, 2, 3), dct=dct)
        times = []
        for _ in range(max(args.steps, 1)):
            torch.cuda.synchronize()
            t = time.perf_counter()
            y, info = eda.decompose(x, axes=(1, 2, 3), return_info=True, dct=dct)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        lib.eda_timer_enable(2)
        _native.check(lib.eda_timer_report(buf, len(buf), err, len(err)), err)
        eda.decompose(x, axes=(1, 2, 3), dct=dct)
        torch.cuda.synchronize()
        _native.check(lib.eda_timer_report(buf, len(buf), err, len(err)), err)
        lib.eda_timer_enable(0)
        d = json.loads(buf.value.decode())["k_dct_fft"]
        us = d["ms"] * 1e3 / max(d["launches"], 1)
        per = d["bytes"] / max(d["launches"], 1)
        flop = (2 * 256 * (3 if dct == "bf16x3" else 1)) * per / 8.0  # 2 n per sample and axis (x3)
        errs = []
        for k in range(y.shape[1]):
            es = [float(torch.linalg.vector_norm((y[b, k] - y64[b, k]).double()) /
                        torch.linalg.vector_norm(y64[b, k].double()).clamp_min(1e-300))
                  for b in range(y.shape[0]) if k <= int(info.n_imfs[b])]
            errs.append(round(max(es), 9) if es else 0.0)
        dt = sorted(times)[len(times) // 2]
        out[dct] = {
            "ms_per_call": round(dt * 1e3, 2),
            "Msamples_per_s": round(n_samples / dt / 1e6, 2),
            "decisions_equal_f64": bool(np.array_equal(info.n_imfs, i64.n_imfs) and
                                        np.array_equal(info.windows, i64.windows)),
            "per_mode_rel_l2_vs_f64": errs,
            "dct_axis_pass": {"launches": d["launches"], "us_per_launch": round(us, 2),
                              "GBs": round(per / (us * 1e-6) / 1e9, 1) if us > 0 else None,
                              "TFLOPs": round(flop / (us * 1e-6) / 1e12, 2) if us > 0 else None},
        }
        log(f"bench: precision {dct}: {dt * 1e3:.1f} ms per call, DCT axis pass {us:.1f} us "
            f"({out[dct]['dct_axis_pass']['GBs']} GB/s), per-mode error vs f64 {errs}")
    return {
        "metric": PREC_METRIC,
        "value": out["bf16x3"]["Msamples_per_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "dtype": "f32 (bf16x3 DCT contractions)",
        "data": "synthetic",
        "config": {"workload": "eda.decompose, 4 x 256 x 256 x 256 float32 plane-wave volumes (3 waves of "
                               "wavelength 6, 20, 64, Gaussian envelope sigma 60, Gaussian trend, N(0, 0.02^2))"},
        "paths": out,
    }


BATCH_METRIC = "Msamples/s decomposed, ragged batch (256 uneven 1-D signals + 32 images, FP32)"


def bench_batching(args, torch) -> dict:
    """The batching workload (256 chirp signals of lengths log-uniform in [1024, 262144] and 32
    images of sides in {128 .. 1024}, float32, device-resident): the per-field form (one
    eda.decompose call per field, in order) against eda.decompose_many (fields grouped by shape,
    the groups on 8 host threads with a stream each). The timed region of each is one pass over
    the workload (median of the timed steps); the rows of both are compared bitwise."""
    import eda
    import workloads

    fs = [torch.from_numpy(f).to("cuda:0") for f in workloads.ragged_fields()]
    n_samples = sum(f.numel() for f in fs)
    torch.cuda.synchronize()
    log(f"bench: batching workload, {len(fs)} fields, {n_samples / 1e6:.1f} Msamples")

    def per_field():
        return [eda.decompose(f) for f in fs]

    def many():
        return eda.decompose_many(fs, workers=8)

    out = {}
    rows = {}
    for name, fn in (("per_field", per_field), ("decompose_many", many)):
        for _ in range(args.warmup):
            fn()
        times = []
        for _ in range(max(args.steps, 1)):
            torch.cuda.synchronize()
            t = time.perf_counter()
            rows[name] = fn()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
        dt = sorted(times)[len(times) // 2]
        out[name] = {"s_per_pass": round(dt, 3), "Msamples_per_s": round(n_samples / dt / 1e6, 2)}
        log(f"bench: batching {name}: {dt:.2f} s per pass, {n_samples / dt / 1e6:.2f} Msamples/s")
    same = all(torch.equal(a, b) for a, b in zip(rows["per_field"], rows["decompose_many"]))
    speedup = out["per_field"]["s_per_pass"] / out["decompose_many"]["s_per_pass"]
    return {
        "metric": BATCH_METRIC,
        "value": out["decompose_many"]["Msamples_per_s"],
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "eda.decompose_many over 256 float32 chirp + FM + quadratic-trend signals (lengths "
                               "log-uniform in [1024, 262144]) and 32 float32 ring / stripe / blob images (sides "
                               "in {128, 256, 384, 512, 768, 1024}), device-resident"},
        "paths": out,
        "speedup_vs_per_field": round(speedup, 2),
        "rows_equal_per_field": bool(same),
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

        threads, _ = host_threads()
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


LEDGER_STEPS = 2
BWD_STEPS = 3


def bench_fwd_bwd(args, x, torch, eda):
    """fwd+bwd: eda.decompose with its tape kept, the differentiation workload's loss gradient
    (sum_k w_k ||row_k - target_k||^2, w_k ~ U(0.5, 2), targets ~ N(0, 1) from a seed) and the
    library's adjoint back to the input; BWD_STEPS timed steps after one warm-up, each bracketed by
    a device synchronization. Returns the rate and the split between the forward and the backward."""
    import numpy as np

    rng = np.random.default_rng(5)

    def step():
        t0 = time.perf_counter()
        td = eda.TapedDecomposition(x)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        P = td.rows.shape[1]
        w = torch.from_numpy(rng.uniform(0.5, 2.0, P)).to(x.device, x.dtype).reshape(1, P, 1, 1)
        tgt = torch.randn(td.rows.shape[1:], generator=gen, device=x.device, dtype=x.dtype)[None]
        g = 2.0 * w * (td.rows - tgt)
        gx = td.backward(g)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        del td, g, gx
        return t1 - t0, t2 - t1

    gen = torch.Generator(device=x.device)
    gen.manual_seed(5)
    step()
    fwd, bwd = [], []
    for _ in range(BWD_STEPS):
        a, b = step()
        fwd.append(a)
        bwd.append(b)
    tot = [a + b for a, b in zip(fwd, bwd)]
    dt = sum(tot) / len(tot)
    log(f"bench: fwd+bwd: {BWD_STEPS} steps, {dt * 1e3:.1f} ms per step (forward with tape "
        f"{sum(fwd) / len(fwd) * 1e3:.1f} ms, loss gradient and backward {sum(bwd) / len(bwd) * 1e3:.1f} ms)")
    return {
        "value": round(B * N * N / dt / 1e6, 3),
        "unit": "Msamples/s",
        "ms_per_step": round(dt * 1e3, 3),
        "fwd_ms": round(sum(fwd) / len(fwd) * 1e3, 3),
        "bwd_ms": round(sum(bwd) / len(bwd) * 1e3, 3),
        "steps": BWD_STEPS,
        "workload": "eda.differentiable forward (tape kept) + loss sum_k w_k ||row_k - target_k||^2 + the "
                    "hand-written adjoint back to the 64 x 512 x 512 float32 input (orientation differentiated)",
    }



def bench_per_item(x, torch, eda, batched_s):
    """The plain per-item form beside the batched one: the same 64 images decomposed one call per
    image (each image's rows bitwise its batched rows), timed over one pass after one untimed pass,
    and the batched call's speed-up over it."""
    for i in range(x.shape[0]):
        eda.decompose(x[i])
    torch.cuda.synchronize()
    t = time.perf_counter()
    for i in range(x.shape[0]):
        eda.decompose(x[i])
    torch.cuda.synchronize()
    per = time.perf_counter() - t
    log(f"bench: per-item form: {per * 1e3:.1f} ms for {x.shape[0]} single-image calls; batched "
        f"{batched_s * 1e3:.1f} ms ({per / batched_s:.2f}x)")
    return {"ms_per_pass": round(per * 1e3, 2), "Msamples_per_s": round(x.numel() / per / 1e6, 2),
            "batched_speedup": round(per / batched_s, 2)}


def kernel_ledger(args, x, lib, torch, eda, _native, step_s):
    """Per-kernel live figures from LEDGER_STEPS extra steps run with every kernel family timed
    (HIP events on the library's stream around each launch; these steps are not the timed ones):
    per kernel its launches per step, average duration, algorithmic bytes per launch, GB/s and
    fraction of the HBM peak, and its share of an un-instrumented step. Also the same figures per
    family: the vHGW pass (k_vhgw_rows + k_vhgw_stream + k_vhgw + both k_box launches + k_cube2)."""
    err = _native.errbuf()
    buf = ctypes.create_string_buffer(4096)
    lib.eda_timer_enable(15)
    _native.check(lib.eda_timer_report(buf, len(buf), err, len(err)), err)  # clears what came before
    for _ in range(LEDGER_STEPS):
        y = eda.decompose(x, axes=(1, 2))
        del y
    torch.cuda.synchronize()
    _native.check(lib.eda_timer_report(buf, len(buf), err, len(err)), err)
    lib.eda_timer_enable(0)
    raw = json.loads(buf.value.decode())

    def figures(ms, launches, nbytes):
        if launches == 0 or ms <= 0:
            return None
        us = ms * 1e3 / launches
        per = nbytes / launches
        gbs = per / (us * 1e-6) / 1e9
        return {
            "launches_per_step": round(launches / LEDGER_STEPS, 1),
            "us_per_launch": round(us, 2),
            "bytes_per_launch": round(per),
            "GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4),
            "share_of_step": round(ms * 1e-3 / LEDGER_STEPS / step_s, 4),
        }

    kernels = {}
    for name, v in raw.items():
        f = figures(v["ms"], v["launches"], v["bytes"])
        if f is not None:
            kernels[name] = f
            log(f"bench: {name:14s} {f['launches_per_step']:7.1f} launches/step {f['us_per_launch']:9.2f} us "
                f"{f['bytes_per_launch'] / 1e6:9.1f} MB {f['GBs']:8.1f} GB/s ({f['frac']:.3f} of peak), "
                f"{100 * f['share_of_step']:.1f} % of a step")
    vh = [raw[k] for k in ("k_vhgw_rows", "k_vhgw", "k_vhgw_stream", "k_box_short", "k_box_long", "k_cube2")
          if k in raw]
    fam = {"vhgw": figures(sum(v["ms"] for v in vh), sum(v["launches"] for v in vh), sum(v["bytes"] for v in vh))}
    return kernels, fam


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-bwd", action="store_true", help="skip the fwd+bwd measurement")
    ap.add_argument("--no-per-item", action="store_true", help="skip the per-item (one call per image) leg")
    ap.add_argument("--fwd-bwd-only", action="store_true",
                    help="run only the fwd+bwd leg (one warm-up step and the timed steps) and print its figures")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one copy kernel of known size (for scaling rocprofv3 byte counters)")
    ap.add_argument("--monogenic", action="store_true",
                    help="measure eda.monogenic on the same batch instead of eda.decompose")
    ap.add_argument("--precision", action="store_true",
                    help="measure the 4 x 256^3 precision workload: exact vs split-BF16 DCT contractions")
    ap.add_argument("--batching", action="store_true",
                    help="measure the ragged batching workload: per-field calls vs eda.decompose_many")
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
    if args.precision:
        print(json.dumps(bench_precision(args, lib, torch)), flush=True)
        return 0
    if args.batching:
        print(json.dumps(bench_batching(args, torch)), flush=True)
        return 0
    log("bench: generating 64 x 512 x 512 float32 inputs")
    x = torch.from_numpy(workloads.rings_stripes_blobs(B, N, np.float32)).to("cuda:0")
    torch.cuda.synchronize()
    if args.monogenic:
        print(json.dumps(bench_monogenic(args, x, lib, torch)), flush=True)
        return 0
    if args.fwd_bwd_only:
        print(json.dumps(bench_fwd_bwd(args, x, torch, eda)), flush=True)
        return 0

    info = None
    warm = []
    for i in range(args.warmup):
        t = time.perf_counter()
        y, info = eda.decompose(x, axes=(1, 2), return_info=True)
        torch.cuda.synchronize()
        warm.append(time.perf_counter() - t)
        log(f"bench: warm-up {i}: {warm[-1]:.3f} s, n_imfs {sorted(set(info.n_imfs.tolist()))}")
        del y

    # the timed steps: no instrumentation (the kernel timer is off)
    lib.eda_timer_enable(0)
    per_step = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        t = time.perf_counter()
        y, info = eda.decompose(x, axes=(1, 2), return_info=True)
        torch.cuda.synchronize()
        per_step.append(time.perf_counter() - t)
        del y
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    med = sorted(per_step)[len(per_step) // 2] if per_step else dt
    log(f"bench: {args.steps} steps, {dt * 1e3:.1f} ms per step (median {med * 1e3:.1f} ms)")

    # the per-kernel ledger: separate, untimed steps with every kernel family timed by HIP events
    kernels, fam = kernel_ledger(args, x, lib, torch, eda, _native, dt)

    if args.copy_probe:
        copy_probe(torch, log)

    # the host-input rate (numpy in, numpy out: PCIe transfers included); stderr only, median of 3
    xh = x.cpu().numpy()
    host = []
    for _ in range(3):
        t = time.perf_counter()
        yh = eda.decompose(xh, axes=(1, 2))
        host.append(time.perf_counter() - t)
        del yh
    host_dt = sorted(host)[1]
    log(f"bench: host input, PCIe-inclusive: {B * N * N / host_dt / 1e6:.1f} Msamples/s "
        f"(median of 3 calls {host_dt * 1e3:.1f} ms: " + ", ".join(f"{h * 1e3:.1f}" for h in host) + " ms)")

    value = B * N * N / dt / 1e6
    vh = fam["vhgw"]
    roofline = {
        "bound": "hbm",
        "kernel": "vHGW pass (k_vhgw_rows, k_vhgw, k_vhgw_stream, k_box, k_cube2)",
        "achieved": vh["GBs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": vh["frac"],
        "traffic": traffic_from_profiles(),
        "algorithmic_bytes_per_launch": vh["bytes_per_launch"],
        "avg_launch_us": vh["us_per_launch"],
        "launches_per_step": vh["launches_per_step"],
        "share_of_step": vh["share_of_step"],
        "kernels": kernels,
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
            "median_ms_per_step": round(med * 1e3, 3),
            "warmup_ms": [round(w * 1e3, 1) for w in warm],
            "host_input_ms": round(host_dt * 1e3, 1),
            "n_imfs": sorted(set(info.n_imfs.tolist())) if info is not None else None,
        },
        "roofline": roofline,
    }
    if not args.no_per_item:
        out["per_item"] = bench_per_item(x, torch, eda, dt)
    if not args.no_bwd:
        out["fwd_bwd"] = bench_fwd_bwd(args, x, torch, eda)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
