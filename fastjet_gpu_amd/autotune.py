"""Autotuner of the multiplicity-bucket table (the launch shape per event size).

For each multiplicity interval (lo, hi] it times every candidate kernel
configuration on a saturating batch of events of that interval alone, on the
current GPU, and keeps the fastest:

  dense engine  every compiled (threads, slots, four-momenta in LDS/global)
                configuration whose capacity threads*slots is at least hi
                (and at most 16*hi, larger ones only waste LDS and lanes);
  tiled engine  tile side R, 1.25 R and 1.5 R, with each compiled kernel
                variant (tile tables of 512, 384, 320 or 256 tiles, and so 8,
                10, 12 or 12 events per CU).

Adjacent intervals with the same winner are merged; the last row is always a
tiled row covering every larger event.

A second, whole-batch stage then times complete tables on a mixed batch (by
default the 4096-event N ~ U[50, 2000] workload): each tiled variant alone, the
merged per-interval table and simpler ones (the per-interval dense winners up
to a cut, then the best tiled variant), and keeps the fastest.  Per-interval throughput does not decide alone:
in a mixed batch the kernels of all buckets share the GPU, the batch ends with
its largest events, and every distinct kernel competes for instruction cache.
The objective of that stage (--objective) is the throughput of a stream of
batches, two batch plans alternating on two streams so that one batch's tail
overlaps the next one's start (pipelined, the default), or of one batch at a
time (single, which favours the variants whose largest events finish first).

The reference picks its launch shapes per call from the occupancy calculator
instead (launch.h:11-60, grid.cu:597-633).

Events of an interval come from the seeded generator: events of the 4096-event
N ~ U[50, 2000] workload whose N falls in the interval, or, outside [50, 2000],
random subsets of the right size of lognormal-multiplicity events.

  python -m fastjet_gpu_amd.autotune [--events 4096] [--objective pipelined|single]
                                     [--json out.json] [--write]

--write replaces csrc/tune_table.inc (the library's default table; rebuild with
python -m fastjet_gpu_amd.build afterwards).  Needs a GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
DEFAULT_EDGES = [32, 64, 96, 128, 192, 256, 384, 512, 768, 1024, 1536, 2048, 4096, 20000]
TILE_QUARTERS = [4, 5, 6]
LAST_ROW = 1000000000


def _sizes(g, workload, count):
    return np.array([g.fjmi_gen_event_size(workload, e) for e in range(count)], np.int64)


class EventSource:
    """Seeded events of a given multiplicity interval (host arrays)."""

    def __init__(self, fj, pool=16384):
        self.fj = fj
        g = fj.load_lib("libfjmi_gen.so")
        self.s3 = _sizes(g, 3, pool)
        self.s5 = _sizes(g, 5, 4096)
        self.cache = {}

    def _event(self, workload, e):
        key = (workload, e)
        if key not in self.cache:
            self.cache[key] = self.fj.gen_event(workload, e)
        return self.cache[key]

    def interval(self, lo, hi, count, distinct=512, seed=1):
        """`count` events (drawn with repetition from `distinct` ones) with lo < N <= hi."""
        rng = np.random.default_rng(seed + 7919 * hi)
        uniq = []
        inside = np.nonzero((self.s3 > lo) & (self.s3 <= hi))[0]
        if lo >= 49 and hi <= 2000 and len(inside) >= 16:
            for e in inside[:distinct]:
                uniq.append(self._event(3, int(e)))
        else:
            for _ in range(distinct):
                n = int(rng.integers(lo + 1, hi + 1))
                src = np.nonzero(self.s5 >= n)[0]
                ev = self._event(5, int(rng.choice(src))) if len(src) else self._event(4, int(rng.integers(0, 64)))
                n = min(n, len(ev))
                uniq.append(ev[np.sort(rng.choice(len(ev), n, replace=False))])
        pick = [uniq[k % len(uniq)] for k in range(count)]
        offsets = np.zeros(count + 1, np.int64)
        np.cumsum([len(e) for e in pick], out=offsets[1:])
        return np.concatenate(pick), offsets


def candidates(configs, lo, hi):
    out = []
    for (cap, eng, T, K, G) in configs:
        if eng != 1 or cap < hi or cap > max(16 * hi, 64) or hi > 8192:
            continue
        out.append((T, K, G))
    rows = [(1, T, K, G) for (T, K, G) in sorted(set(out))]
    variants = sorted({G for (cap, eng, T, K, G) in configs if eng == 2})
    rows += [(2, 64, q, v) for v in variants for q in TILE_QUARTERS]
    return rows


def time_config(fj, torch, eng, cfg, hi, p4, off, d, p, R, reps):
    engine_id, T, K, G = cfg
    if engine_id == 1:
        eng.set_buckets([(hi, 1, T, K, G), (LAST_ROW, 2, 64, 0, 0)])
    else:
        eng.set_buckets([(LAST_ROW, 2, 64, K, G)])
    b = eng.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), None, None, p, R, d["h1"].data_ptr(),
                         d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
    try:
        t = time.perf_counter()
        b.launch()
        b.sync()
        once = time.perf_counter() - t
        inner = max(1, int(0.02 / max(once, 1e-6)))  # each sample spans >= ~20 ms of GPU work
        best = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            for _ in range(inner):
                b.launch()
            b.sync()
            best = min(best, (time.perf_counter() - t) / inner)
        return (len(off) - 1) / best
    finally:
        b.close()


def merge(choices):
    """[(hi, cfg)] -> bucket rows, merging neighbours with the same cfg; last row tiled, open-ended."""
    rows = []
    for hi, (e, T, K, G) in choices:
        row = [hi, e, T, K, G]
        if rows and rows[-1][1:] == row[1:]:
            rows[-1][0] = hi
        else:
            rows.append(row)
    if rows[-1][1] == 2:
        rows[-1][0] = LAST_ROW
    else:
        tiled = [c for _, c in choices if c[0] == 2]
        q, v = (tiled[-1][2], tiled[-1][3]) if tiled else (0, 0)
        rows.append([LAST_ROW, 2, 64, q, v])
    return [tuple(r) for r in rows]


def tiled_alone(configs):
    """One tiled row for every event, per compiled kernel variant (tile side R)."""
    variants = sorted({G for (cap, eng, T, K, G) in configs if eng == 2})
    return [(f"tiled variant {v} alone", [(LAST_ROW, 2, 64, 4, v)]) for v in variants]


def simple_tables(choices, last):
    """Dense per-interval winners up to a cut, then the tiled row `last` (engine, threads, slots, variant)."""
    out = []
    for cut in (128, 256, 512):
        dense = [(hi, c) for hi, c in choices if hi <= cut and c[0] == 1]
        if not dense or dense[-1][0] != cut:
            continue
        rows = [r for r in merge(dense + [(LAST_ROW, last)])]
        out.append((f"dense<= {cut} + tiled", rows))
        if dense:
            one = (cut, dense[-1][1])
            out.append((f"one dense row <= {cut} + tiled", merge([one, (LAST_ROW, last)])))
    return out


def time_table(eng, rows, p4, off, d, p, R, reps):
    """Events/s of one batch launch at a time (best of the repetitions)."""
    eng.set_buckets(rows)
    b = eng.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), None, None, p, R, d["h1"].data_ptr(),
                         d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
    try:
        b.launch()
        b.sync()
        best = float("inf")
        for _ in range(max(reps, 5)):
            t = time.perf_counter()
            b.launch()
            b.sync()
            best = min(best, time.perf_counter() - t)
        return (len(off) - 1) / best
    finally:
        b.close()


def time_table_pipelined(torch, eng, rows, off, ds, streams, p, R, reps, launches=12):
    """Events/s of a stream of batches: one plan per output set in `ds`, launched in turn on `streams`, so that
    a batch's tail overlaps the next batch's start (best of the repetitions of `launches` launches)."""
    eng.set_buckets(rows)
    plans = [eng.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), None, None, p, R, d["h1"].data_ptr(),
                              d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
             for d in ds]
    try:
        for b, s in zip(plans, streams):
            b.launch(s.cuda_stream)
            b.sync(s.cuda_stream)
        best = float("inf")
        for _ in range(max(reps, 3)):
            torch.cuda.synchronize()
            t = time.perf_counter()
            for k in range(launches):
                plans[k % len(plans)].launch(streams[k % len(plans)].cuda_stream)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t) / launches)
        for b, s in zip(plans, streams):
            b.sync(s.cuda_stream)
        return (len(off) - 1) / best
    finally:
        for b in plans:
            b.close()


def device_buffers(torch, dev, p4, off):
    tot = len(p4)
    return dict(p4=torch.from_numpy(p4).to(dev), off=torch.from_numpy(off).to(dev),
                h1=torch.empty(tot, dtype=torch.int32, device=dev), h2=torch.empty(tot, dtype=torch.int32, device=dev),
                hd=torch.empty(tot, dtype=torch.float64, device=dev),
                j=torch.empty(4 * tot, dtype=torch.float64, device=dev),
                nj=torch.empty(len(off) - 1, dtype=torch.int32, device=dev))


def coop_top_row(rows):
    """The table with its last row (every larger event) on 8 waves per event (threads 512): a plan runs
    that row cooperatively while its bucket holds at most one event per CU (large events, few of them)
    and one wave per event otherwise, so the whole-batch stage's timings, whose largest events are in
    full buckets or absent, hold for it too."""
    rows = [tuple(r) for r in rows]
    last = rows[-1]
    if last[1] == 2:
        rows[-1] = (last[0], 2, 512, last[3], last[4])
    return rows


def write_inc(rows, note):
    rows = coop_top_row(rows)
    path = os.path.join(PKG, "csrc", "tune_table.inc")
    lines = ["/* Multiplicity buckets: {max_n, engine, threads, slots, four-momenta",
             " * in global memory} (engine 1 = dense; engine 2 = tiled, slots = tile side",
             " * in quarters of R with 0 = R, and the last field the kernel variant);",
             " * rows sorted by max_n.",
             f" * Written by fastjet_gpu_amd/autotune.py: {note} */"]
    lines += ["{%d, %d, %d, %d, %d}," % r for r in rows]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--events", type=int, default=4096, help="events per interval and candidate")
    ap.add_argument("--edges", default=",".join(map(str, DEFAULT_EDGES)))
    ap.add_argument("--p", type=int, default=-1)
    ap.add_argument("--R", type=float, default=0.4)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--batch-workload", type=int, default=3, help="generator workload of the whole-batch stage")
    ap.add_argument("--batch-events", type=int, default=4096)
    ap.add_argument("--objective", choices=("pipelined", "single"), default="pipelined",
                    help="whole-batch stage: a stream of batches on two alternating plans, or one batch at a time")
    ap.add_argument("--json", default=None)
    ap.add_argument("--write", action="store_true")
    a = ap.parse_args(argv)

    sys.path.insert(0, os.path.dirname(PKG))
    import torch
    import fastjet_gpu_amd as fj

    eng = fj.Engine(0)
    saved = eng.buckets()
    configs = eng.kernel_configs()
    src = EventSource(fj)
    edges = [int(x) for x in a.edges.split(",")]
    dev = torch.device("cuda:0")
    results, choices = [], []
    lo = 0
    try:
        for hi in edges:
            count = a.events if hi <= 2048 else max(512, min(a.events, int(8e6 // ((lo + hi) / 2))))
            p4, off = src.interval(lo, hi, count)
            d = device_buffers(torch, dev, p4, off)
            best = None
            for cfg in candidates(configs, lo, hi):
                rate = time_config(fj, torch, eng, cfg, hi, p4, off, d, a.p, a.R, a.reps)
                results.append(dict(lo=lo, hi=hi, engine=cfg[0], threads=cfg[1], slots=cfg[2], p4_global=cfg[3],
                                    events_per_s=rate, mean_n=float(np.mean(np.diff(off)))))
                if best is None or rate > best[0]:
                    best = (rate, cfg)
            print(f"({lo:6d},{hi:6d}]  best {best[1]}  {best[0]:.0f} ev/s", flush=True)
            choices.append((hi, best[1]))
            lo = hi
            del d
        # whole-batch stage
        p4, off = fj.gen_workload(a.batch_workload, a.batch_events)
        d = device_buffers(torch, dev, p4, off)
        if a.objective == "pipelined":
            d2 = device_buffers(torch, dev, p4, off)
            d2["p4"], d2["off"] = d["p4"], d["off"]
            streams = [torch.cuda.current_stream(), torch.cuda.Stream()]

            def timed(rows):
                return time_table_pipelined(torch, eng, rows, off, [d, d2], streams, a.p, a.R, a.reps)
        else:
            def timed(rows):
                return time_table(eng, rows, p4, off, d, a.p, a.R, a.reps)
        whole = []
        alone = tiled_alone(configs)
        for name, rows in alone:
            rate = timed(rows)
            whole.append(dict(name=name, table=[list(r) for r in rows], events_per_s=rate))
            print(f"batch of {len(off) - 1} (workload {a.batch_workload}): {name:32s} {rate:10.0f} ev/s  {rows}",
                  flush=True)
        last = tuple(max(whole, key=lambda w: w["events_per_s"])["table"][0][1:])
        tables = [("per-interval winners", merge(choices))] + simple_tables(choices, last)
        if len(choices) > 1 and choices[-1][1][0] == 2 and choices[-1][1] != last:
            # the largest events keep their own winner (the whole batch rarely has them)
            tables.append((f"best tiled, largest interval's above {choices[-2][0]}",
                           merge([(choices[-2][0], last), (LAST_ROW, choices[-1][1])])))
        tables.append(("library default", saved))
        for name, rows in tables:
            rate = timed(rows)
            whole.append(dict(name=name, table=[list(r) for r in rows], events_per_s=rate))
            print(f"batch of {len(off) - 1} (workload {a.batch_workload}): {name:32s} {rate:10.0f} ev/s  {rows}",
                  flush=True)
        del d
        if a.objective == "pipelined":
            del d2
    finally:
        eng.set_buckets(saved)
    best_whole = max(whole, key=lambda w: w["events_per_s"])
    rows = [tuple(r) for r in best_whole["table"]]
    props = torch.cuda.get_device_properties(0)
    dev_name = f"{props.name} ({getattr(props, 'gcnArchName', '?')})"
    out = dict(device=dev_name, p=a.p, R=a.R, events=a.events, edges=edges, objective=a.objective, measurements=results,
               whole_batch=whole, chosen=best_whole["name"], table=[list(r) for r in rows])
    print(json.dumps(dict(table=out["table"])), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(out, fh, indent=1)
    if a.write:
        note = (f"whole-batch stage ({a.objective}) on {a.batch_events} events of workload {a.batch_workload}, "
                f"p={a.p} R={a.R}: {best_whole['name']}")
        print("wrote", write_inc(rows, f"{note}, {dev_name}"))
    eng.close()


if __name__ == "__main__":
    main()
