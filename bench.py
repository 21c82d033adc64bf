#!/usr/bin/env python3
"""Benchmark of the MI355X jet-clustering engine (one GPU).

Workload: a batch of 4096 events with N ~ U{50..2000} particles (seeded generator,
workload 3 of fastjet_gpu_amd/csrc/fjmi_gen.c), anti-kt, R = 0.4, FP64.

A step is one pass of the hot path over the batch: ONE hipGraph launch whose
persistent workgroups cluster all 4096 events (inputs already resident in HBM).
`value` is events/s over the K timed steps.  Beside it the line reports:
  plain        the one-event-per-launch path (one kernel launch + sync per event)
               on a sample of the same events, and the batched/plain ratio;
  host_memory_path  the same batch from host arrays (PCIe copies and planning
               included; never `value`);
  latency_ms   p50/p99 per-event latency of the plain path and per-batch latency;
  roofline     the dominant (slowest-finishing) bucket kernel's algorithmic FP64
               rate (dense-equivalent operation count) against MI355X's FP64 peak;
  cpu_baseline the CPU oracle (tiled N2Plain, FastJet N2Tiled-style) on the host cores.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "events/s batched vs per-launch at N~U[50,2000], anti-kt R=0.4; p99 latency"
FP64_PEAK_TFLOPS = 78.6  # MI355X published FP64 peak (vector and matrix)


def measured_traffic(kernel_tag):
    """HBM bytes per launch of the dominant kernel on this workload, from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/), or None."""
    path = os.path.join(ROOT, "profiles", "tiled_kernel_traffic_workload3.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:") and kernel_tag in line:
                return int(line.split()[1])
    except OSError:
        pass
    return None


def flops(sum_n2, sum_n):
    """Algorithmic FP64 operations of the dense nearest-neighbour method:
    8 per pair evaluation for N(N-1)/2 initial pairs plus N(N-1)/2 merged-row
    refreshes, plus N(N+1)/2 argmin compares:  8 N(N-1) + N(N+1)/2 per event."""
    return 8.0 * (sum_n2 - sum_n) + 0.5 * (sum_n2 + sum_n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=4096)
    ap.add_argument("--cpu-repeat", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before the timed region, fill and copy a 256 MiB buffer: reference kernels of known HBM "
                         "traffic for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (events are independent, so several GPUs would each run "
              "this same batch on their own share of the events; that is not run here). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import torch
    import fastjet_gpu_amd as fj

    p, R = fj.ANTIKT, 0.4
    p4, off = fj.gen_workload(3, a.events)
    ne = len(off) - 1
    sizes = np.diff(off)
    sha = hashlib.sha256(p4.tobytes()).hexdigest()

    dev = torch.device("cuda:0")
    eng = fj.Engine(0)
    tot = len(p4)
    d_p4 = torch.from_numpy(p4).to(dev)
    d_off = torch.from_numpy(off).to(dev)
    d_h1 = torch.zeros(tot, dtype=torch.int32, device=dev)
    d_h2 = torch.zeros(tot, dtype=torch.int32, device=dev)
    d_hd = torch.zeros(tot, dtype=torch.float64, device=dev)
    d_j = torch.zeros(tot * 4, dtype=torch.float64, device=dev)
    d_nj = torch.zeros(ne, dtype=torch.int32, device=dev)
    if a.calibrate_copy:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        src.fill_(1.0)  # writes 256 MiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst
    stream = torch.cuda.current_stream().cuda_stream
    batch = eng.batch_create(off, d_p4.data_ptr(), d_off.data_ptr(), None, None, p, R, d_h1.data_ptr(),
                             d_h2.data_ptr(), d_hd.data_ptr(), d_j.data_ptr(), d_nj.data_ptr())
    for _ in range(a.warmup):
        batch.launch(stream)
    batch.sync(stream)

    # ---- timed region: K graph launches, inputs resident ----
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        starts[k].record()
        batch.launch(stream)
        ends[k].record()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    batch.sync(stream)
    wall = t1 - t0
    step_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    value = ne * a.steps / wall
    bstats = batch.stats()

    # ---- correctness spot check of the timed outputs against the oracle ----
    parity = "not checked"
    try:
        from oracle import oracle as orc
        h1, h2 = d_h1.cpu().numpy(), d_h2.cpu().numpy()
        hd, jj, nj = d_hd.cpu().numpy(), d_j.cpu().numpy().reshape(-1, 4), d_nj.cpu().numpy()
        ok = 0
        picks = list(range(0, ne, max(1, ne // 8)))[:8]
        for e in picks:
            lo, hi = off[e], off[e + 1]
            w = orc.n2plain(p4[lo:hi], p, R)
            g = (h1[lo:hi], h2[lo:hi], hd[lo:hi], jj[lo:lo + nj[e]])
            ok += all(np.asarray(x).tobytes() == np.asarray(y).tobytes() for x, y in zip(w, g))
        parity = f"bit-exact history and jets on {ok}/{len(picks)} sampled events"
    except Exception as ex:  # the oracle is optional at bench time
        parity = f"not checked ({ex})"

    # ---- plain one-event-per-launch path on a sample ----
    ns = min(a.plain_sample, ne)
    lat = np.zeros(ns)
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    for e in range(ns):
        lo = int(off[e])
        t = time.perf_counter()
        eng.cluster_event_device(d_p4.data_ptr() + 32 * lo, int(sizes[e]), p, R, d_h1.data_ptr() + 4 * lo,
                                 d_h2.data_ptr() + 4 * lo, d_hd.data_ptr() + 8 * lo, d_j.data_ptr() + 32 * lo,
                                 d_nj.data_ptr() + 4 * e, stream)
        torch.cuda.current_stream().synchronize()
        lat[e] = time.perf_counter() - t
    plain_wall = time.perf_counter() - tp0
    plain_value = ns / plain_wall

    # ---- host-memory path (PCIe-inclusive): copy in, plan, one graph launch, copy out ----
    eng.cluster_batch(p4, off, p, R)  # warm
    th0 = time.perf_counter()
    eng.cluster_batch(p4, off, p, R)
    host_path = ne / (time.perf_counter() - th0)

    # ---- roofline of the dominant bucket kernel ----
    dom = max(bstats, key=lambda b: b["ms"])
    dom_flops = flops(dom["sum_n2"], dom["particles"])
    achieved = dom_flops / (dom["ms"] * 1e-3) / 1e12 if dom["ms"] > 0 else None
    all_flops = flops(float(np.sum(sizes.astype(np.float64) ** 2)), float(tot))
    variants = {0: "tiled_kernel<512, 2, false>", 1: "tiled_kernel<384, 3, false>", 2: "tiled_kernel<256, 4, false>",
                3: "tiled_kernel<128, 4, false>"}
    traffic = measured_traffic(variants.get(dom["p4_global"], "?")) if dom["engine"] == 2 and ne == 4096 else None

    # ---- CPU baseline: the oracle on the host cores, bounded sample ----
    cpu = None
    if not a.no_cpu:
        try:
            from oracle import oracle as orc
            cs = min(a.cpu_sample, ne)
            threads = a.cpu_threads or min(16, os.cpu_count() or 1)
            secs = min(orc.time_batch(p4[:off[cs]], off[:cs + 1], p, R, threads, "tiled")
                       for _ in range(max(1, a.cpu_repeat)))
            cs2 = min(cs, 1024)
            secs_plain = orc.time_batch(p4[:off[cs2]], off[:cs2 + 1], p, R, threads, "n2plain")
            cpu = {"value": cs / secs, "unit": "events/s", "cores": threads, "kind": "port",
                   "sample": f"the first {cs} events of the same batch clustered by oracle/'s tiled N2Plain "
                             f"(FastJet N2Tiled-style) on {threads} host threads, best of {max(1, a.cpu_repeat)}: "
                             f"{secs:.2f} s wall; the plain N2Plain restatement did {cs2} events in "
                             f"{secs_plain:.2f} s ({cs2 / secs_plain:.0f} events/s)"}
        except Exception as ex:
            cpu = {"value": None, "unit": "events/s", "cores": 0, "kind": "port", "sample": f"failed: {ex}"}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "events/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic: seeded generator workload 3, sha256 {sha[:16]}",
        "config": {"workload": f"{ne} events, N~U[50,2000] (mean {sizes.mean():.0f}), anti-kt R=0.4, "
                               f"one hipGraph launch of persistent bucket kernels",
                   "events": ne, "particles": int(tot), "algorithm": "antikt", "R": R},
        "plain": {"value": plain_value, "unit": "events/s", "sample_events": ns,
                  "desc": "one kernel launch + stream sync per event, inputs resident"},
        "ratio_batched_vs_plain": value / plain_value,
        "host_memory_path": {"value": host_path, "unit": "events/s",
                             "desc": "fjmi_cluster_batch from host arrays: H2D, plan + graph capture, launch, D2H"},
        "latency_ms": {"plain_p50": float(np.percentile(lat, 50) * 1e3),
                       "plain_p99": float(np.percentile(lat, 99) * 1e3),
                       "batch_p50": float(np.percentile(step_ms, 50)),
                       "batch_p99": float(np.percentile(step_ms, 99))},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": (achieved / FP64_PEAK_TFLOPS) if achieved else None, "traffic": traffic,
                     "kernel": (f"dense_kernel<{dom['threads']}, {dom['slots']}, {str(bool(dom['p4_global'])).lower()}, 1>"
                                if dom["engine"] == 1 else variants.get(dom["p4_global"], "tiled_kernel"))
                               + f" (bucket N<={min(dom['max_n'], int(sizes.max()))}, {dom['n_events']} events, "
                               f"{dom['ms']:.2f} ms)",
                     "fp64_ops_per_launch": dom_flops, "batch_fp64_ops": all_flops,
                     "batch_rate_tflops": all_flops / (np.median(step_ms) * 1e-3) / 1e12},
        "buckets": bstats,
        "parity": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    batch.close()
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
