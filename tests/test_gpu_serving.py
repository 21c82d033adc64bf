"""Serving (fjmi_server_*): events admitted under a memory budget, clustered by
persistent GPU workgroups as they arrive, results announced per event.

Bar: every result bit-exact against the oracle (history and jets), and equal to
the same event clustered alone; backpressure, idle relaunch and argument errors
behave as documented in include/fjmi.h.  Every test prints its case first.
"""
import time

import numpy as np
import pytest

from oracle import oracle as orc


def oracle_result(p4, p, R):
    return orc.n2plain(p4, p, R) if len(p4) <= 3000 else orc.tiled(p4, p, R)


def same(want, r):
    h1, h2, hd, jets = want
    return (np.array_equal(h1, r["hist_p1"]) and np.array_equal(h2, r["hist_p2"])
            and hd.tobytes() == r["hist_dij"].tobytes()
            and np.asarray(jets).reshape(-1, 4).tobytes() == r["jets"].tobytes())


def drive(srv, requests, deadline_s=120.0):
    """Submits (tag, p4, p, R) in order, keeping rejected ones queued; returns {tag: result}."""
    pending = list(requests)
    done = {}
    t_end = time.time() + deadline_s
    rejected = 0
    while pending or len(done) < len(requests):
        while pending:
            tag, p4, p, R = pending[0]
            if not srv.submit(p4, p, R, tag):
                rejected += 1
                break
            pending.pop(0)
        for r in srv.poll():
            done[r["tag"]] = r
        assert time.time() < t_end, f"serving stalled: {len(done)}/{len(requests)} done"
    return done, rejected


@pytest.mark.gpu
@pytest.mark.parametrize("variant", [1, 0])
def test_serving_matches_oracle(engine, variant):
    """Mixed sizes (workload 3 and lognormal workload-5 events), per-event algorithm
    and radius, a budget small enough that admission has to wait; the default
    tiled kernel variant (384 tiles) and the 512-tile one."""
    import fastjet_gpu_amd as fj
    reqs = []
    for k in range(160):
        ev = fj.gen_event(3, 2000 + k) if k % 4 else fj.gen_event(5, k)
        p, R, _ = fj.gen_serving_attrs(k)
        reqs.append((k, ev, p, R))
    print(f"[gpu serving] variant {variant}: {len(reqs)} events, sizes {min(len(r[1]) for r in reqs)}.."
          f"{max(len(r[1]) for r in reqs)}", flush=True)
    srv = engine.server(device_budget=48 << 20, host_budget=24 << 20, max_events=48, variant=variant)
    try:
        done, rejected = drive(srv, reqs)
        info = srv.info()
    finally:
        srv.close()
    print(f"[gpu serving] rejected submits {rejected}, info {info}", flush=True)
    assert rejected > 0  # the budget forced waiting
    assert info["completed"] == len(reqs) and info["in_flight"] == 0
    for tag, ev, p, R in reqs:
        r = done[tag]
        assert r["n"] == len(ev) and r["njets"] >= 1
        assert same(oracle_result(ev, p, R), r), f"event {tag} (N={len(ev)}, p={p}, R={R})"


@pytest.mark.gpu
def test_serving_equals_alone_and_batch(engine):
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 24, first=700)
    reqs = [(e, p4[off[e]:off[e + 1]], -1, 0.4) for e in range(24)]
    print("[gpu serving] 24 events vs alone and batch", flush=True)
    srv = engine.server(device_budget=256 << 20, host_budget=64 << 20, max_events=256)
    try:
        done, _ = drive(srv, reqs)
    finally:
        srv.close()
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, -1, 0.4)
    for e, ev, p, R in reqs:
        alone = engine.cluster_event(ev, p, R)
        assert same(alone, done[e]), f"event {e} alone"
        a, b = off[e], off[e + 1]
        assert same((h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]]), done[e]), f"event {e} batch"


@pytest.mark.gpu
def test_serving_backpressure_and_empty_event(engine):
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(1, 3)
    print("[gpu serving] slots full -> submit refused; empty event", flush=True)
    srv = engine.server(device_budget=64 << 20, host_budget=32 << 20, max_events=4)
    try:
        for k in range(4):
            assert srv.submit(ev, -1, 0.4, k)
        assert not srv.submit(ev, -1, 0.4, 99)  # all slots in flight
        got = {}
        t_end = time.time() + 60
        while len(got) < 4:
            for r in srv.poll():
                got[r["tag"]] = r
            assert time.time() < t_end
        assert srv.submit(np.zeros((0, 4)), 1, 0.4, 7)
        t_end = time.time() + 60
        res = []
        while not res:
            res = srv.poll()
            assert time.time() < t_end
        assert res[0]["tag"] == 7 and res[0]["n"] == 0 and res[0]["njets"] == 0
        want = orc.n2plain(ev, -1, 0.4)
        assert all(same(want, got[k]) for k in range(4))
    finally:
        srv.close()


@pytest.mark.gpu
def test_serving_relaunch_after_idle(engine):
    """The GPU workgroups leave after idle_us without work; the next submit relaunches."""
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(2, 1)
    want = orc.n2plain(ev, 1, 0.4)
    print("[gpu serving] idle 2 ms, three submits 50 ms apart", flush=True)
    srv = engine.server(device_budget=64 << 20, host_budget=32 << 20, max_events=8, idle_us=2000)
    try:
        for k in range(3):
            assert srv.submit(ev, 1, 0.4, k)
            t_end = time.time() + 60
            res = []
            while not res:
                res = srv.poll()
                assert time.time() < t_end
            assert same(want, res[0])
            time.sleep(0.05)
        assert srv.info()["launches"] >= 3
    finally:
        srv.close()


@pytest.mark.gpu
def test_server_argument_errors(engine):
    import fastjet_gpu_amd as fj
    print("[gpu serving] argument errors", flush=True)
    with pytest.raises(fj.FjmiError):
        engine.server(variant=9)
    with pytest.raises(fj.FjmiError):
        engine.server(tile_q=3)
    with pytest.raises(fj.FjmiError):
        engine.server(device_budget=1 << 20, host_budget=1 << 20, max_events=1024)
    srv = engine.server(device_budget=16 << 20, host_budget=4 << 20, max_events=16)
    try:
        ev = fj.gen_event(1, 0)
        with pytest.raises(fj.FjmiError):
            srv.submit(ev, 2, 0.4, 0)
        with pytest.raises(fj.FjmiError):
            srv.submit(ev, 1, -0.4, 0)
        cap = srv.info()["particle_capacity"]
        with pytest.raises(fj.FjmiError):
            srv.submit(np.ones((cap + 1, 4)), 1, 0.4, 0)
    finally:
        srv.close()


@pytest.mark.gpu
def test_serve_replays_an_event_file(tmp_path):
    """fjmi_serve --replay: the load generator serves the events of a binary event file with their own
    exponent and radius (every result is checked for a GPU failure by fjmi_serve itself)."""
    import json
    import os
    import subprocess
    import fastjet_gpu_amd as fj
    print("case: fjmi_serve --replay, 64 events, 200 requests at 2000/s", flush=True)
    p4, off = fj.gen_workload(5, 64)
    ne = len(off) - 1
    path = str(tmp_path / "replay.fjmi")
    fj.write_events(path, p4, off, np.array([(-1, 0, 1)[e % 3] for e in range(ne)], np.int32),
                    np.array([(0.4, 0.8)[e % 2] for e in range(ne)]))
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "fastjet_gpu_amd", "bin",
                       "fjmi_serve")
    r = subprocess.run([exe, "--replay", path, "--rates", "2000", "--events", "200", "--device-gib", "2",
                        "--host-gib", "1"], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
    assert len(lines) == 1 and lines[0]["requests"] == 200 and lines[0]["unique_events"] == ne
    assert abs(lines[0]["mean_n_unique"] - float(np.mean(np.diff(off)))) < 0.06
