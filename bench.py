#!/usr/bin/env python3
"""Benchmark of the MI355X jet-clustering engine (one GPU).

Workload: a batch of 4096 events with N ~ U{50..2000} particles (seeded generator,
workload 3 of fastjet_gpu_amd/csrc/fjmi_gen.c), anti-kt, R = 0.4, FP64.

A step is one pass of the hot path over the batch: ONE hipGraph launch whose
persistent workgroups cluster all 4096 events (inputs already resident in HBM).
`value` is events/s over the K timed steps, with two batch plans (each with its
own outputs) alternating on two streams, as a batch pipeline would run them: the
tail of one launch, where only its largest events still run, overlaps the start
of the next.  Beside it the line reports:
  single_stream  the same K launches on one stream (each batch alone);
  plain        the one-event-per-launch path (one kernel launch + sync per event)
               on a sample of the same events, and the batched/plain ratio;
  launch_floor the reference's tri_matrix_empty schedule (2N+3 empty launches
               per event, tri_matrix_empty.cu:257-292) on a sample of events;
  host_memory_path  the same batch from host arrays (PCIe copies included; never
               `value`);
  latency_ms   p50/p99 per-event latency of the plain path and per-batch latency;
  serving      the continuous-batching server on the serving stream (lognormal
               multiplicities, mixed algorithms and radii) at two Poisson rates:
               served events/s and per-event p50/p99 latency;
  configs      the single-event configurations (N = 354; N = 808 under three
               algorithms and two radii): the GPU's plain per-event time beside
               the CPU oracle on one thread; the high-pileup batch on the GPU
               (device-resident, and through the host-memory path);
  roofline     the tiled kernel against MI355X peaks: algorithmic FP64 work of
               the tiled method (pair distances counted by the oracle) per
               launch, the dense-equivalent count, and rocprofv3 counters
               (executed FP64 FLOPs, VALU busy, LDS and HBM bytes) from the
               committed profile of the same kernel source;
  cpu_baseline the CPU oracle (tiled N2Plain, FastJet N2Tiled-style) on all the
               host's cores, best of 3; its results also check every event of
               the timed batch bit for bit (`parity`).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
       python bench.py --summarize DIR [DIR ...]   (rocprofv3 output -> profile summary)
       python bench.py --highpileup                (the high-pileup batch alone, one JSON line)
"""
import argparse
import csv
import glob
import hashlib
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "events/s batched vs per-launch at N~U[50,2000], anti-kt R=0.4; p99 latency"
# MI355X (gfx950) published peaks and clock
FP64_PEAK_TFLOPS = 78.6          # vector FP64, FMA counted as 2
HBM_PEAK_GBPS = 8000.0
N_CU, N_SIMD, SCLK_GHZ = 256, 1024, 2.4
LDS_PEAK_GBPS = N_CU * 128 * SCLK_GHZ  # 128 bytes per clock per CU
COUNTER_PROFILE = os.path.join(ROOT, "profiles", "tiled_kernel_counters_workload3.md")
# the tiled kernel of each variant (the p4_global field of a tiled bucket row)
TILED_VARIANTS = {0: "tiled_kernel<512, 2, false>", 1: "tiled_kernel<384, 3, false>", 2: "tiled_kernel<320, 3, false>",
                  3: "tiled_kernel<256, 3, false>"}
PROFILED_KERNEL = TILED_VARIANTS[3]  # the default table's kernel for the 4096-event batch
KERNEL_SOURCES = ["fastjet_gpu_amd/csrc/tiled.hip", "fastjet_gpu_amd/csrc/tiled_body.inc", "fastjet_gpu_amd/csrc/fjmi_math.h",
                  "fastjet_gpu_amd/csrc/fjmi_math_tables.h", "fastjet_gpu_amd/csrc/fjmi_internal.h",
                  "fastjet_gpu_amd/csrc/tune_table.inc", "fastjet_gpu_amd/csrc/runtime.hip", "fastjet_gpu_amd/build.py"]


def source_sha():
    """Hash of the sources that define the tiled kernel and its launch on this workload."""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def dense_equivalent_flops(sum_n2, sum_n):
    """FP64 operations of the dense nearest-neighbour method: 8 per pair
    evaluation for N(N-1)/2 initial pairs plus N(N-1)/2 merged-row refreshes,
    plus N(N+1)/2 argmin compares: 8 N(N-1) + N(N+1)/2 per event."""
    return 8.0 * (sum_n2 - sum_n) + 0.5 * (sum_n2 + sum_n)


# ---------------------------------------------------------------------------
# rocprofv3 output -> committed summary (bench.py --summarize), and back
# ---------------------------------------------------------------------------
def _read_csvs(dirs, name):
    rows = []
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", f"*{name}"), recursive=True):
            with open(path, newline="") as fh:
                rows += list(csv.DictReader(fh))
    return rows


# the rocprofv3 runs the committed counter profile comes from (one pass each)
PROFILE_BENCH = "python3 bench.py --steps 4 --warmup 1 --no-extras --no-cpu --pipeline 1"
PMC_PASSES = [
    ("A", "SQ_INSTS_VALU_FLOPS_FP64 SQ_INSTS_VALU_FMA_F64 SQ_INSTS_VALU_ADD_F64 SQ_INSTS_VALU_MUL_F64 "
          "SQ_INSTS_VALU_TRANS_F64 SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_VALU_INT64", ""),
    ("B", "SQ_INSTS_LDS_LOAD_BANDWIDTH SQ_INSTS_LDS_STORE_BANDWIDTH SQ_INSTS_LDS_ATOMIC_BANDWIDTH SQ_INSTS_LDS "
          "SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_VALU SQ_BUSY_CU_CYCLES SQ_ACTIVE_INST_LDS", ""),
    ("C", "SQ_WAVE_CYCLES SQ_WAIT_INST_ANY SQ_WAIT_ANY SQ_ACTIVE_INST_ANY SQ_WAVES SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR "
          "SQ_ACTIVE_INST_SCA", ""),
    ("D", "FETCH_SIZE", "--calibrate-copy"),
    ("E", "WRITE_SIZE", "--calibrate-copy"),
    ("F", "TCC_HIT_sum TCC_MISS_sum", ""),
]


def profile_commands(out="gpurun_out/prof"):
    cmds = [f"timeout -k 10 300 rocprofv3 --kernel-trace --stats -d {out}/kt -o run --output-format csv -- "
            f"{PROFILE_BENCH}"]
    for tag, ctrs, extra in PMC_PASSES:
        cmds.append(f"timeout -s KILL 180 rocprofv3 --pmc {ctrs} -d {out}/{tag} -o run --output-format csv -- "
                    f"{PROFILE_BENCH} {extra}".rstrip())
    cmds.append(f"python3 bench.py --summarize {out}/kt " + " ".join(f"{out}/{t}" for t, _, _ in PMC_PASSES))
    return cmds


def summarize(dirs, kernel_key=PROFILED_KERNEL):
    """Per-launch counter values of the tiled kernel (and the calibration copy
    kernel's FETCH_SIZE / WRITE_SIZE) from rocprofv3 counter_collection.csv files,
    and its average duration from kernel_stats.csv, as machine-readable lines."""
    per = {}   # (kernel, dispatch) -> {counter: value}
    for r in _read_csvs(dirs, "counter_collection.csv"):
        k = (r.get("Kernel_Name", ""), r.get("Dispatch_Id", ""))
        per.setdefault(k, {})
        c = r["Counter_Name"]
        per[k][c] = per[k].get(c, 0.0) + float(r["Counter_Value"])
    want = {}
    copy = {}
    for (name, _), vals in per.items():
        short = name.replace("fjmi::", "")
        if kernel_key in short:
            for c, v in vals.items():
                want.setdefault(c, []).append(v)
        elif "copy" in name.lower():
            # the 256 MiB device-to-device copy is the largest copy dispatch of the run
            for c, v in vals.items():
                copy[c] = [max(copy.get(c, [0.0])[0], v)]
    out = [f"source_sha: {source_sha()}", f"kernel: {kernel_key}"]
    steps = 4146882  # clustering steps per launch = particles of the 4096-event batch
    for r in _read_csvs(dirs, "kernel_stats.csv"):
        if kernel_key in r.get("Name", "").replace("fjmi::", ""):
            out.append(f"kernel_avg_ms: {float(r['AverageNs']) / 1e6:.4f} (calls {r.get('Calls', '?')})")
    for c, vs in sorted(want.items()):
        out.append(f"counter {c}: {np.mean(vs):.6g} (dispatches {len(vs)})")
    for c, vs in sorted(copy.items()):
        out.append(f"copy_counter {c}: {np.mean(vs):.6g} (256 MiB read and 256 MiB written by the copy kernel)")
    # HBM bytes per launch: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each scaled by
    # what the same counter reads for the copy kernel of known traffic (256 MiB each way)
    if "FETCH_SIZE" in want and "WRITE_SIZE" in want and "FETCH_SIZE" in copy and "WRITE_SIZE" in copy:
        ff = (256 << 20) / (np.mean(copy["FETCH_SIZE"]) * 1024.0)
        wf = (256 << 20) / (np.mean(copy["WRITE_SIZE"]) * 1024.0)
        fb = np.mean(want["FETCH_SIZE"]) * 1024.0 * ff
        wb = np.mean(want["WRITE_SIZE"]) * 1024.0 * wf
        out.append(f"fetch_factor: {ff:.4f}")
        out.append(f"write_factor: {wf:.4f}")
        out.append(f"fetch_bytes_per_launch: {fb:.6g}")
        out.append(f"write_bytes_per_launch: {wb:.6g}")
        out.append(f"traffic_bytes_per_launch: {fb + wb:.6g}")
    g = {c: float(np.mean(v)) for c, v in want.items()}
    rows = []
    for label, c in (("VALU instructions", "SQ_INSTS_VALU"), ("  of which FP64 add/sub", "SQ_INSTS_VALU_ADD_F64"),
                     ("  FP64 mul", "SQ_INSTS_VALU_MUL_F64"), ("  FP64 fma", "SQ_INSTS_VALU_FMA_F64"),
                     ("  FP64 transcendental (rcp)", "SQ_INSTS_VALU_TRANS_F64"), ("  64-bit integer", "SQ_INSTS_VALU_INT64"),
                     ("SALU instructions", "SQ_INSTS_SALU"), ("LDS instructions", "SQ_INSTS_LDS"),
                     ("vector memory reads", "SQ_INSTS_VMEM_RD"), ("vector memory writes", "SQ_INSTS_VMEM_WR")):
        if c in g:
            rows.append(f"| {label} | {g[c]:.4g} | {g[c] / steps:.1f} |")
    extra = []
    if "SQ_WAVE_CYCLES" in g:
        for label, c in (("waiting (SQ_WAIT_ANY)", "SQ_WAIT_ANY"), ("issuing (SQ_ACTIVE_INST_ANY)", "SQ_ACTIVE_INST_ANY"),
                         ("ready, not issued (SQ_WAIT_INST_ANY)", "SQ_WAIT_INST_ANY")):
            if c in g:
                extra.append(f"| share of wave cycles {label} | {g[c] / g['SQ_WAVE_CYCLES']:.3f} |")
    if "TCC_HIT_sum" in g and "TCC_MISS_sum" in g:
        extra.append(f"| L2 hit rate | {g['TCC_HIT_sum'] / (g['TCC_HIT_sum'] + g['TCC_MISS_sum']):.3f} |")
    doc = [f"# {kernel_key} counters, 4096-event batch (N ~ U[50, 2000], anti-kt R = 0.4)", "",
           "Commands (repository root on one MI355X, after `cd /tmp && export TMPDIR=/tmp`), one pass each:", "",
           "```", *profile_commands(), "```", "",
           "Machine-readable lines (read by bench.py; counter values per launch, averaged over the profiled "
           "launches; `source_sha` hashes the kernel's sources as profiled, and bench.py derives no fraction "
           "from counters of other sources):", "", "```", *out, "```", "",
           f"Per launch and per clustering step ({steps} steps per launch, one per particle):", "",
           "| counter | per launch | per step |", "|---|---|---|", *rows, "", "| ratio | value |", "|---|---|",
           *extra, ""]
    print("\n".join(doc))
    return 0


def read_profile(path=COUNTER_PROFILE):
    """Counters per launch from the committed summary: (dict, sha, kernel_avg_ms, kernel name)."""
    vals, sha, avg, kern = {}, None, None, None
    try:
        for line in open(path):
            line = line.strip().strip("`")
            if line.startswith("source_sha:"):
                sha = line.split()[1]
            elif line.startswith("kernel:"):
                kern = line.split(":", 1)[1].strip()
            elif line.startswith("kernel_avg_ms:"):
                avg = float(line.split()[1])
            elif line.startswith("counter "):
                name, v = line[len("counter "):].split(":", 1)
                vals[name.strip()] = float(v.split()[0])
            elif line.startswith("traffic_bytes_per_launch:"):
                vals["traffic_bytes_per_launch"] = float(line.split()[1])
    except OSError:
        pass
    return vals, sha, avg, kern


def counter_fields(t_s, n_steps_total, kname):
    """Fractions of MI355X peaks from the committed counters at kernel time t_s, if they were taken on the same
    kernel (`kname`) built from the same sources."""
    vals, sha, avg, kern = read_profile()
    cur = source_sha()
    if not vals:
        return {"source": None, "stale": None}
    stale = sha != cur or kern != kname
    f = {"source": os.path.relpath(COUNTER_PROFILE, ROOT), "profiled_source_sha": sha, "current_source_sha": cur,
         "profiled_kernel": kern, "stale": stale, "profiled_kernel_ms": avg}
    if stale:  # counters of another kernel or build: not combined with this run's time
        return f
    fl = vals.get("SQ_INSTS_VALU_FLOPS_FP64")  # per wave instruction (FMA = 2), not per lane
    if fl is not None:
        f["fp64_flops_executed"] = fl * 64.0  # all 64 lanes counted: an upper bound of the lane work
        f["fp64_tflops_executed"] = fl * 64.0 / t_s / 1e12
        f["fp64_issue_frac"] = fl * 64.0 / t_s / 1e12 / FP64_PEAK_TFLOPS
        f["fp64_insts_per_step"] = sum(vals.get(c, 0.0) for c in (
            "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64", "SQ_INSTS_VALU_TRANS_F64")
        ) / n_steps_total if n_steps_total else None
    av = vals.get("SQ_ACTIVE_INST_VALU")  # quad-cycles, summed over waves
    if av is not None:
        f["valu_busy_frac"] = av * 4.0 / (t_s * SCLK_GHZ * 1e9 * N_SIMD)
    iv = vals.get("SQ_INSTS_VALU")
    isa = vals.get("SQ_INSTS_SALU")
    if iv is not None and n_steps_total:
        f["valu_insts_per_step"] = iv / n_steps_total
    if isa is not None and n_steps_total:
        f["salu_insts_per_step"] = isa / n_steps_total
    lds = sum(vals.get(c, 0.0) for c in ("SQ_INSTS_LDS_LOAD_BANDWIDTH", "SQ_INSTS_LDS_STORE_BANDWIDTH",
                                         "SQ_INSTS_LDS_ATOMIC_BANDWIDTH")) * 64.0
    if lds:
        f["lds_bytes"] = lds
        f["lds_gbps"] = lds / t_s / 1e9
        f["lds_peak_gbps"] = LDS_PEAK_GBPS
        f["lds_frac"] = lds / t_s / 1e9 / LDS_PEAK_GBPS
    tr = vals.get("traffic_bytes_per_launch")
    if tr:
        f["hbm_gbps"] = tr / t_s / 1e9
        f["hbm_peak_gbps"] = HBM_PEAK_GBPS
        f["hbm_frac"] = tr / t_s / 1e9 / HBM_PEAK_GBPS
    wc = vals.get("SQ_WAVE_CYCLES")
    if wc:
        for key, c in (("wait_frac", "SQ_WAIT_ANY"), ("wait_issue_frac", "SQ_WAIT_INST_ANY"),
                       ("issuing_frac", "SQ_ACTIVE_INST_ANY")):
            if vals.get(c) is not None:
                f[key] = vals[c] / wc
    h, m = vals.get("TCC_HIT_sum"), vals.get("TCC_MISS_sum")
    if h is not None and m:
        f["l2_hit_frac"] = h / (h + m)
    return f


# ---------------------------------------------------------------------------
def host_info():
    cores = len(os.sched_getaffinity(0))
    model = "?"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def timed_plain(eng, d, ev_off, sizes, idx, p, R, stream, engine=0):
    import torch
    lat = np.zeros(len(idx))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, e in enumerate(idx):
        lo = int(ev_off[e])
        t = time.perf_counter()
        eng.cluster_event_device(d["p4"].data_ptr() + 32 * lo, int(sizes[e]), p, R, d["h1"].data_ptr() + 4 * lo,
                                 d["h2"].data_ptr() + 4 * lo, d["hd"].data_ptr() + 8 * lo,
                                 d["j"].data_ptr() + 32 * lo, d["nj"].data_ptr() + 4 * e, stream, engine=engine)
        torch.cuda.current_stream().synchronize()
        lat[k] = time.perf_counter() - t
    return time.perf_counter() - t0, lat


def highpileup_device(eng, fj, torch, p, R, launches=3):
    """The high-pileup batch (64 events, N ~ U[5000, 20000]) from device memory, one graph launch per
    batch on the default table (its top row: 8 waves per event while the bucket holds at most one event per
    CU): ms per launch (mean of `launches` after one warm-up launch), the history digest and the bucket
    statistics."""
    hp4, hoff = fj.gen_workload(4, 64)
    tot, ne = len(hp4), len(hoff) - 1
    dev = torch.device("cuda:0")
    d = dict(p4=torch.from_numpy(hp4).to(dev), off=torch.from_numpy(hoff).to(dev),
             h1=torch.zeros(tot, dtype=torch.int32, device=dev), h2=torch.zeros(tot, dtype=torch.int32, device=dev),
             hd=torch.zeros(tot, dtype=torch.float64, device=dev), j=torch.zeros(4 * tot, dtype=torch.float64, device=dev),
             nj=torch.zeros(ne, dtype=torch.int32, device=dev))
    b = eng.batch_create(hoff, d["p4"].data_ptr(), d["off"].data_ptr(), None, None, p, R, d["h1"].data_ptr(),
                         d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
    try:
        b.launch()
        b.sync()
        t = time.perf_counter()
        for _ in range(launches):
            b.launch()
        b.sync()
        ms = (time.perf_counter() - t) / launches * 1e3
        stats = b.stats()
    finally:
        b.close()
    digest = hashlib.sha256(d["h1"].cpu().numpy().tobytes() + d["hd"].cpu().numpy().tobytes()).hexdigest()[:16]
    return ms, digest, stats


def serving_leg(rates, seconds):
    exe = os.path.join(ROOT, "fastjet_gpu_amd", "bin", "fjmi_serve")
    if not os.path.exists(exe):
        return {"error": "bin/fjmi_serve not built"}
    r = subprocess.run([exe, "--rates", ",".join(f"{x:g}" for x in rates), "--seconds", str(seconds),
                        "--unique", "1024"], capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        return {"error": r.stderr[-500:]}
    out = []
    for line in r.stdout.splitlines():
        if line.startswith("{"):
            j = json.loads(line)
            out.append({"offered_per_s": j["rate"], "events_per_s": j["events_per_s"],
                        "mid_run_events_per_s": j["mid_events_per_s"], "requests": j["requests"],
                        "p50_ms": j["latency_ms"]["p50"], "p99_ms": j["latency_ms"]["p99"],
                        "refused_submits": j["refused_submits"]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--events", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--floor-sample", type=int, default=32)
    ap.add_argument("--cpu-repeat", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core this process may run on")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--pipeline", type=int, default=2,
                    help="batch plans alternating on as many streams (each with its own outputs), so that one "
                         "batch's tail overlaps the next batch's start; 1: one stream")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the timed batch (for profiling runs): no plain, floor, host, serving, configs legs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="before the timed region, fill and copy a 256 MiB buffer: reference kernels of known HBM "
                         "traffic for scaling rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--summarize", nargs="+", metavar="DIR", help="summarise rocprofv3 output directories and exit")
    ap.add_argument("--highpileup", action="store_true",
                    help="time only the high-pileup batch (64 events, N ~ U[5000, 20000], R = 0.4, kt and anti-kt) "
                         "from device memory and print one JSON line")
    ap.add_argument("--events-file", metavar="FILE",
                    help="time the events of a binary event file (fjmi_events_write) instead of the generated "
                         "workload, with the file's per-event exponent and radius where it holds them")
    a = ap.parse_args()
    if a.summarize:
        return summarize(a.summarize)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: measures one GPU only (events are independent, so several GPUs would each run "
              "this same batch on their own share of the events; that is not run here). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import torch
    import fastjet_gpu_amd as fj

    if a.highpileup:
        eng = fj.Engine(0)
        out = {"workload": "high-pileup batch: 64 events, N ~ U[5000, 20000], R = 0.4, device-resident, "
                           "one graph launch per batch, default table"}
        for pp, name in ((fj.KT, "kt"), (fj.ANTIKT, "antikt")):
            ms, digest, stats = highpileup_device(eng, fj, torch, pp, 0.4)
            out[name] = {"ms_per_batch": ms, "events_per_s": 64e3 / ms, "history_sha": digest,
                         "buckets": [{k: st[k] for k in ("max_n", "threads", "p4_global", "n_events", "grid")}
                                     for st in stats]}
        eng.close()
        print(json.dumps(out), flush=True)
        return 0

    p, R = fj.ANTIKT, 0.4
    pa = Ra = None
    if a.events_file:
        p4, off, pa, Ra = fj.read_events(a.events_file)
    else:
        p4, off = fj.gen_workload(3, a.events)
    ne = len(off) - 1
    if ne < 1:
        print("bench.py: no events to time", file=sys.stderr)
        return 2
    sizes = np.diff(off)
    sha = hashlib.sha256(p4.tobytes()).hexdigest()

    dev = torch.device("cuda:0")
    eng = fj.Engine(0)
    tot = len(p4)
    d = dict(p4=torch.from_numpy(p4).to(dev), off=torch.from_numpy(off).to(dev),
             h1=torch.zeros(tot, dtype=torch.int32, device=dev), h2=torch.zeros(tot, dtype=torch.int32, device=dev),
             hd=torch.zeros(tot, dtype=torch.float64, device=dev), j=torch.zeros(tot * 4, dtype=torch.float64, device=dev),
             nj=torch.zeros(ne, dtype=torch.int32, device=dev))
    d["p"] = torch.from_numpy(pa).to(dev) if pa is not None else None
    d["R"] = torch.from_numpy(Ra).to(dev) if Ra is not None else None
    dptr = lambda t: t.data_ptr() if t is not None else None
    if a.calibrate_copy:
        src = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        src.fill_(1.0)  # writes 256 MiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)  # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del src, dst
    stream = torch.cuda.current_stream().cuda_stream
    batch = eng.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), dptr(d["p"]), dptr(d["R"]), p, R,
                             d["h1"].data_ptr(),
                             d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
    for _ in range(a.warmup):
        batch.launch(stream)
    batch.sync(stream)

    # extra plans (own outputs, own stream) for --pipeline > 1
    plans = [(batch, torch.cuda.current_stream(), d)]
    for _ in range(a.pipeline - 1):
        dk = dict(d)
        for key in ("h1", "h2", "hd", "j", "nj"):
            dk[key] = torch.zeros_like(d[key])
        bk = eng.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), dptr(d["p"]), dptr(d["R"]), p, R,
                              dk["h1"].data_ptr(),
                              dk["h2"].data_ptr(), dk["hd"].data_ptr(), dk["j"].data_ptr(), dk["nj"].data_ptr())
        sk = torch.cuda.Stream()
        bk.launch(sk.cuda_stream)
        bk.sync(sk.cuda_stream)
        plans.append((bk, sk, dk))

    # ---- one stream: each launch alone on the GPU (the batch latency and the kernel time) ----
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(a.steps)]
    torch.cuda.synchronize()
    ts0 = time.perf_counter()
    for k in range(a.steps):
        starts[k].record()
        batch.launch(stream)
        ends[k].record()
    torch.cuda.synchronize()
    single_wall = time.perf_counter() - ts0
    batch.sync(stream)
    step_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    bstats = batch.stats()

    # ---- timed region: K graph launches, the plans alternating on their streams ----
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        bk, sk, _ = plans[k % len(plans)]
        bk.launch(sk.cuda_stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for bk, sk, _ in plans:
        bk.sync(sk.cuda_stream)
    wall = t1 - t0
    value = ne * a.steps / wall
    print(f"[bench] timed {a.steps} launches: {value:.0f} events/s", file=sys.stderr, flush=True)

    outs = [dict(h1=dk["h1"].cpu().numpy(), h2=dk["h2"].cpu().numpy(), hd=dk["hd"].cpu().numpy(),
                 j=dk["j"].cpu().numpy().reshape(-1, 4), nj=dk["nj"].cpu().numpy()) for _, _, dk in plans]

    # ---- CPU leg: the oracle on every host core (baseline), whose results check every timed event ----
    cores, model = host_info()
    threads = a.cpu_threads or cores
    cpu, parity, pairs = None, "not checked", None
    orc = None
    try:
        from oracle import oracle as orc
    except Exception as ex:  # the oracle is test infrastructure; the bench runs without it
        parity = f"not checked ({ex})"
    if orc is not None:
        runs = []
        for k in range(max(1, a.cpu_repeat) if not a.no_cpu else 1):
            runs.append(orc.cluster_batch(p4, off, p if pa is None else pa, R if Ra is None else Ra, threads,
                                          "tiled"))
        w = runs[0]
        bad = 0
        for g in outs:
            for e in range(ne):
                lo, hi = off[e], off[e + 1]
                if (w[4][e] != g["nj"][e] or w[0][lo:hi].tobytes() != g["h1"][lo:hi].tobytes()
                        or w[1][lo:hi].tobytes() != g["h2"][lo:hi].tobytes()
                        or w[2][lo:hi].tobytes() != g["hd"][lo:hi].tobytes()
                        or w[3][lo:lo + w[4][e]].tobytes() != g["j"][lo:lo + w[4][e]].tobytes()):
                    bad += 1
        nchk = ne * len(outs)
        parity = (f"bit-exact history and jets on {nchk - bad}/{nchk} events (every event of the timed batch, "
                  f"in each of the {len(outs)} plans' outputs)")
        pairs = float(np.sum(w[5]))
        if not a.no_cpu:
            secs = min(r[6] for r in runs)
            cpu = {"value": ne / secs, "unit": "events/s", "cores": threads, "kind": "port",
                   "host_cores": cores, "host_cpu": model,
                   "sample": f"all {ne} events of the timed batch clustered by oracle/'s tiled N2Plain "
                             f"(FastJet N2Tiled-style, g++ -O3) on {threads} threads of {cores} host cores, "
                             f"best of {len(runs)}: {secs:.2f} s wall"}
        print(f"[bench] CPU leg done: {parity}", file=sys.stderr, flush=True)

    extras = {}
    if not a.no_extras and not a.events_file:
        # ---- plain one-event-per-launch path on a sample ----
        ns = min(a.plain_sample, ne)
        plain_wall, lat = timed_plain(eng, d, off, sizes, range(ns), p, R, stream)
        eng.stream_check(stream)
        extras["plain"] = {"value": ns / plain_wall, "unit": "events/s", "sample_events": ns,
                           "desc": "one kernel launch + stream sync per event, inputs resident"}
        extras["latency_ms"] = {"plain_p50": float(np.percentile(lat, 50) * 1e3),
                                "plain_p99": float(np.percentile(lat, 99) * 1e3),
                                "batch_p50": float(np.percentile(step_ms, 50)),
                                "batch_p99": float(np.percentile(step_ms, 99))}

        # ---- the reference's per-step launch floor: 2N+3 empty launches per event ----
        nf = min(a.floor_sample, ne)
        rec = torch.zeros(40 * int(sizes[:nf].max()), dtype=torch.uint8, device=dev)
        torch.cuda.synchronize()
        tf0 = time.perf_counter()
        for e in range(nf):
            eng.cluster_pseudojets(rec.data_ptr(), int(sizes[e]), p, R, engine=fj.ENGINE_EMPTY, stream=stream)
            torch.cuda.current_stream().synchronize()
        floor_wall = time.perf_counter() - tf0
        launches = int(np.sum(2 * sizes[:nf] + 3))
        extras["launch_floor"] = {"value": nf / floor_wall, "unit": "events/s", "sample_events": nf,
                                  "us_per_launch": floor_wall / launches * 1e6,
                                  "desc": "tri_matrix_empty schedule: 2N+3 empty kernel launches per event, "
                                          "one sync per event (no clustering)"}

        # ---- host-memory path (PCIe-inclusive) ----
        eng.cluster_batch(p4, off, p, R)  # warm: the context's staging grows once
        hts = []
        for _ in range(3):
            th0 = time.perf_counter()
            eng.cluster_batch(p4, off, p, R)
            hts.append(time.perf_counter() - th0)
        extras["host_memory_path"] = {"value": ne / float(np.median(hts)), "unit": "events/s", "calls": len(hts),
                                      "desc": "fjmi_cluster_batch from pageable host arrays (median of 3 calls): "
                                              "staging copies, H2D, kernels, D2H of the history and the packed "
                                              "jets, in four chunks on four streams"}

        # ---- single-event configurations and the high-pileup batch ----
        cfg = {}
        ev1 = fj.gen_event(1, 0)
        ev2 = fj.gen_event(2, 0)
        runs2 = [(pp, RR) for pp in (1, 0, -1) for RR in (0.4, 1.0)]

        def gpu_event_ms(ev, pp, RR, reps=20):
            n = len(ev)
            dd = dict(p4=torch.from_numpy(np.ascontiguousarray(ev)).to(dev),
                      h1=torch.zeros(n, dtype=torch.int32, device=dev), h2=torch.zeros(n, dtype=torch.int32, device=dev),
                      hd=torch.zeros(n, dtype=torch.float64, device=dev),
                      j=torch.zeros(4 * n, dtype=torch.float64, device=dev), nj=torch.zeros(1, dtype=torch.int32, device=dev))
            one = np.array([0, n], np.int64)
            timed_plain(eng, dd, one, np.array([n]), [0], pp, RR, stream)
            _, lt = timed_plain(eng, dd, one, np.array([n]), [0] * reps, pp, RR, stream)
            return float(np.mean(lt) * 1e3)

        cfg["N354_antikt_0.4"] = {"gpu_plain_ms": gpu_event_ms(ev1, -1, 0.4)}
        for pp, RR in runs2:
            cfg[f"N808_{ {1: 'kt', 0: 'cam', -1: 'antikt'}[pp]}_{RR}"] = {"gpu_plain_ms": gpu_event_ms(ev2, pp, RR)}
        if orc is not None:
            def cpu_event_ms(ev, pp, RR, reps):
                evs = np.tile(ev, (reps, 1))
                o = np.arange(reps + 1, dtype=np.int64) * len(ev)
                return orc.cluster_batch(evs, o, pp, RR, 1, "n2plain")[6] / reps * 1e3
            cfg["N354_antikt_0.4"]["cpu_1thread_ms"] = cpu_event_ms(ev1, -1, 0.4, 400)
            for pp, RR in runs2:
                cfg[f"N808_{ {1: 'kt', 0: 'cam', -1: 'antikt'}[pp]}_{RR}"]["cpu_1thread_ms"] = cpu_event_ms(ev2, pp, RR, 100)
        hp4, hoff = fj.gen_workload(4, 64)
        for pp, name in ((1, "kt"), (-1, "antikt")):
            ms, _, hst = highpileup_device(eng, fj, torch, pp, 0.4)
            eng.cluster_batch(hp4, hoff, pp, 0.4)
            th = time.perf_counter()
            eng.cluster_batch(hp4, hoff, pp, 0.4)
            cfg[f"highpileup64_{name}_0.4"] = {
                "gpu_batch_ms": ms, "gpu_batch_events_per_s": 64e3 / ms,
                "waves_per_event": max(st["threads"] for st in hst) // 64,
                "gpu_batch_events_per_s_host_path": 64 / (time.perf_counter() - th)}
        extras["configs"] = cfg
        print("[bench] plain, floor, host path and configs done", file=sys.stderr, flush=True)

    # ---- roofline of the dominant kernel ----
    dom = max(bstats, key=lambda b: b["ms"])
    t_kernel = float(np.mean(step_ms)) * 1e-3  # HIP events on the launch stream around each graph launch
    kname = (f"dense_kernel<{dom['threads']}, {dom['slots']}, {str(bool(dom['p4_global'])).lower()}, 1>"
             if dom["engine"] == 1 else TILED_VARIANTS.get(dom["p4_global"], "tiled_kernel"))
    dense_f = dense_equivalent_flops(float(np.sum(sizes.astype(np.float64) ** 2)), float(tot))
    alg_f = 8.0 * pairs if pairs is not None else None
    ctr = (counter_fields(t_kernel, float(tot), kname) if (len(bstats) == 1 and ne == 4096 and not a.events_file)
           else {"source": None})
    achieved = alg_f / t_kernel / 1e12 if alg_f else None
    roof = {"bound": "latency",
            "limiter": "the per-step dependent chain of one wave per event (argmin, merge, gather, neighbour "
                       "repair), times the events in flight (workgroups per CU); one batch alone ends with its "
                       "largest events, which a stream of batches overlaps with the next batch",
            "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS if achieved else None,
            "traffic": None,
            "kernel": f"{kname} (bucket N<={min(dom['max_n'], int(sizes.max()))}, {dom['n_events']} events, "
                      f"{dom['grid']} workgroups)",
            "kernel_ms": t_kernel * 1e3, "kernel_ms_device_clock": dom["ms"],
            "algorithmic": {"pair_evals": pairs, "fp64_ops": alg_f,
                            "per_unit": "8 FP64 operations per pair distance; pair distances of the tiled method "
                                        "counted by the oracle's tiled restatement on the same events"},
            "dense_equivalent": {"fp64_ops": dense_f, "tflops": dense_f / t_kernel / 1e12,
                                 "frac": dense_f / t_kernel / 1e12 / FP64_PEAK_TFLOPS,
                                 "per_unit": "8 N(N-1) + N(N+1)/2 per event (dense nearest-neighbour method)"},
            "counters": ctr}
    vals, psha, _, _ = read_profile()
    if vals.get("traffic_bytes_per_launch") and not ctr.get("stale", True):
        roof["traffic"] = vals["traffic_bytes_per_launch"]

    serving = None
    if not a.no_extras and not a.events_file and a.serve_seconds > 0:  # own process, after the legs above
        serving = {"stream": "lognormal N (median 1500, sigma 0.6, cap 20000); {ak 0.4, ak 0.8, CA 0.8, kt 0.4}; "
                             "16 GiB device budget", "rates": serving_leg([1e4, 1e5, 1e6], a.serve_seconds)}
        print("[bench] serving leg done", file=sys.stderr, flush=True)

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
    }
    if a.events_file:
        line["data"] = f"event file {os.path.basename(a.events_file)}, sha256 of four-momenta {sha[:16]}"
        line["config"] = {"workload": f"{ne} events of {os.path.basename(a.events_file)} (mean N "
                                      f"{sizes.mean():.0f}, max {int(sizes.max())}), "
                                      + ("per-event p" if pa is not None else "anti-kt") + ", "
                                      + ("per-event R" if Ra is not None else "R=0.4")
                                      + ", one hipGraph launch of persistent bucket kernels",
                          "events": ne, "particles": int(tot)}
    line["single_stream"] = {"value": ne * a.steps / single_wall, "unit": "events/s",
                             "ms_per_step": single_wall * 1e3 / a.steps,
                             "desc": "the same K launches on one stream: each batch alone on the GPU"}
    line["pipeline"] = {"plans": len(plans), "desc": "value: batch plans (own outputs) alternating on "
                                                     f"{len(plans)} streams, so a batch's tail overlaps the next "
                                                     "batch's start; every step clusters all events"}
    if "plain" in extras:
        line["plain"] = extras["plain"]
        line["ratio_batched_vs_plain"] = value / extras["plain"]["value"]
    for k in ("launch_floor", "host_memory_path", "latency_ms"):
        if k in extras:
            line[k] = extras[k]
    line["roofline"] = roof
    line["buckets"] = bstats
    line["parity"] = parity
    line["cpu_baseline"] = cpu
    if serving is not None:
        line["serving"] = serving
    if "configs" in extras:
        line["configs"] = extras["configs"]
    print(json.dumps(line), flush=True)
    batch.close()
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
