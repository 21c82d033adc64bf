"""GPU parity: the HIP dense engine against the CPU oracle, through the C ABI.

Bar: bit-exact merge history (parents and dij, compared as raw bits) and
bit-exact jets, on the reference's own input files (tests/golden/*.dat), on the
seeded workloads and on edge cases.  Every test prints the case it runs before
touching the GPU.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import GOLDEN

ALGOS = [(1, "kt"), (0, "cam"), (-1, "antikt")]


def same(a, b):
    """Bitwise equality of two oracle/engine results (hist_p1, hist_p2, hist_dij, jets)."""
    h1, h2, hd, j = a
    g1, g2, gd, gj = b
    return (np.array_equal(h1, g1) and np.array_equal(h2, g2) and hd.tobytes() == gd.tobytes()
            and np.asarray(j).tobytes() == np.asarray(gj).tobytes())


def first_diff(a, b):
    for s in range(len(a[0])):
        if a[0][s] != b[0][s] or a[1][s] != b[1][s] or a[2][s].tobytes() != b[2][s].tobytes():
            return s, (a[0][s], a[1][s], a[2][s]), (b[0][s], b[1][s], b[2][s])
    return None


ENGINES = [(1, "dense"), (2, "tiled")]


def check_event(engine, p4, p, R, label, eng=0):
    print(f"[gpu parity] {label} N={len(p4)} p={p} R={R} engine={eng}", flush=True)
    want = orc.n2plain(p4, p, R) if len(p4) <= 6000 else orc.tiled(p4, p, R)
    got = engine.cluster_event(p4, p, R, engine=eng)
    assert same(want, got), f"{label}: first differing step {first_diff(want, got)}"


@pytest.mark.gpu
@pytest.mark.parametrize("eng,ename", ENGINES)
@pytest.mark.parametrize("fname", ["354_fastjet.dat", "808_cms.dat"])
@pytest.mark.parametrize("p,alg", ALGOS)
@pytest.mark.parametrize("R", [0.4, 1.0])
def test_reference_files(engine, fname, p, alg, R, eng, ename):
    ev = orc.read_events(os.path.join(GOLDEN, fname))[0]
    check_event(engine, ev, p, R, f"{fname}/{alg}/{ename}", eng)


@pytest.mark.gpu
@pytest.mark.parametrize("eng,ename", ENGINES)
def test_multi_event_file(engine, eng, ename):
    events = orc.read_events(os.path.join(GOLDEN, "multi-event.dat"))
    assert [len(e) for e in events] == [354, 713, 676, 552, 520, 294, 582, 515, 417, 837]
    for k, ev in enumerate(events):
        for p, alg in ALGOS:
            check_event(engine, ev, p, 0.4, f"multi-event[{k}]/{alg}/{ename}", eng)


@pytest.mark.gpu
@pytest.mark.parametrize("eng,ename", ENGINES)
@pytest.mark.parametrize("workload", [1, 2])
@pytest.mark.parametrize("p,alg", ALGOS)
def test_generated_single(engine, workload, p, alg, eng, ename):
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(workload, 0)
    for R in (0.4, 1.0, 3.0):
        check_event(engine, ev, p, R, f"workload{workload}/{alg}/{ename}", eng)


@pytest.mark.gpu
@pytest.mark.parametrize("p,alg", [(1, "kt"), (-1, "antikt")])
def test_high_multiplicity_tiled(engine, p, alg):
    """Workload 4 (N ~ U[5000, 20000]) on the tiled engine, R = 0.4."""
    import fastjet_gpu_amd as fj
    for e in (0, 2):
        ev = fj.gen_event(4, e)
        check_event(engine, ev, p, 0.4, f"workload4[{e}]/{alg}/tiled", 2)


@pytest.mark.gpu
@pytest.mark.parametrize("p,alg", [(1, "kt"), (-1, "antikt")])
def test_high_pileup_batch(engine, p, alg):
    """The first 24 events of workload 4 (N ~ U[5000, 20000]) in one batch, each
    bit-exact against the oracle: gathers of several chunks and their
    incremental tile-minimum refresh (for kt, the merged jet's larger weight can
    raise a neighbour's diJ; event 17 exercises it)."""
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(4, 24)
    print(f"[gpu parity] workload4 batch of 24 events, {len(p4)} particles, {alg}", flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, p, 0.4)
    for e in range(24):
        a, b = off[e], off[e + 1]
        want = orc.tiled(p4[a:b], p, 0.4)
        got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
        assert same(want, got), f"event {e}: {first_diff(want, got)}"


@pytest.mark.gpu
@pytest.mark.parametrize("var", [2, 3])
def test_high_multiplicity_small_tile_table(engine, var):
    """Workload-4 events through a batch on the 320- and 256-tile kernels (tiles
    taller in rapidity than R), kt and anti-kt R = 0.4."""
    import fastjet_gpu_amd as fj
    evs = [fj.gen_event(4, e) for e in (0, 5)]
    off = np.array([0, len(evs[0]), len(evs[0]) + len(evs[1])], np.int64)
    p4 = np.concatenate(evs)
    saved = engine.buckets()
    try:
        engine.set_buckets([(1000000000, 2, 64, 0, var)])
        for p in (1, -1):
            print(f"[gpu parity] workload4 events 0,5 N={off[1]},{off[2] - off[1]} p={p} variant {var}", flush=True)
            h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, p, 0.4)
            for e in range(2):
                a, b = off[e], off[e + 1]
                want = orc.tiled(p4[a:b], p, 0.4)
                got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
                assert same(want, got), f"event {e}: {first_diff(want, got)}"
    finally:
        engine.set_buckets(saved)


@pytest.mark.gpu
def test_every_bucket_size(engine):
    """One event per multiplicity bucket boundary (and just above it), under the
    library's table and under a mixed dense/tiled table."""
    import fastjet_gpu_amd as fj
    rng = np.random.default_rng(7)
    big = fj.gen_event(4, 3)
    saved = engine.buckets()
    mixed = [(64, 1, 64, 1, 0), (256, 1, 64, 4, 0), (1024, 2, 64, 5, 3), (1000000000, 2, 64, 4, 0)]
    try:
        for table in (saved, mixed):
            engine.set_buckets(table)
            for (max_n, eng, threads, slots, g) in table:
                for n in sorted({max(1, max_n - 1), max_n, max_n + 1}):
                    if n > len(big):
                        continue
                    sel = big[rng.permutation(len(big))[:n]]
                    check_event(engine, sel, -1, 0.4, f"bucket<= {max_n} (engine {eng} {threads}x{slots},g{g})")
    finally:
        engine.set_buckets(saved)


@pytest.mark.gpu
@pytest.mark.parametrize("eng,ename", ENGINES)
def test_edge_cases(engine, eng, ename):
    one = np.array([[1.0, 2.0, 3.0, 10.0]])
    check_event(engine, one, 1, 0.4, "single particle", eng)
    two = np.array([[1.0, 0.0, 0.0, 1.0], [1.0, 0.01, 0.0, 1.0001]])
    for p, alg in ALGOS:
        check_event(engine, two, p, 0.4, f"two particles/{alg}", eng)
    # exact duplicates: many exact distance ties and orphans
    base = np.array([[3.0, 1.0, 0.5, 4.0], [0.5, 2.0, -1.0, 3.0], [-1.0, -1.0, 0.2, 2.0]])
    dup = np.repeat(base, 40, axis=0)
    for p, alg in ALGOS:
        check_event(engine, dup, p, 0.7, f"duplicates/{alg}", eng)
    # zero-pt and beam-like momenta (rapidity +-(1e5+|pz|))
    zp = np.array([[0.0, 0.0, 5.0, 5.0], [0.0, 0.0, -7.0, 7.0], [1.0, 1.0, 0.0, 2.0], [0.0, 0.0, 1.0, 2.0],
                   [1e-3, 0.0, 0.0, 1e-3], [0.0, 0.0, 0.0, 0.0]])
    for p, alg in ALGOS:
        check_event(engine, zp, p, 0.4, f"zero pt/{alg}", eng)
    # a regular lattice: many exact geometric ties
    g = []
    for i in range(12):
        for j in range(12):
            eta, phi = -1.0 + 0.25 * i, 0.25 * j
            pt = 1.0 + 0.125 * ((i * 7 + j * 3) % 5)
            g.append([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)])
    for p, alg in ALGOS:
        check_event(engine, np.array(g), p, 0.3, f"lattice/{alg}", eng)


@pytest.mark.gpu
@pytest.mark.parametrize("eng", [1, 2])
def test_empty_event(engine, eng):
    h1, h2, hd, jets = engine.cluster_event(np.zeros((0, 4)), -1, 0.4, engine=eng)
    assert len(h1) == 0 and len(jets) == 0


@pytest.mark.gpu
@pytest.mark.parametrize("tq,var", [(0, 0), (6, 1), (16, 2), (5, 3), (4, 3)])
def test_batch_all_tiled(engine, tq, var):
    """A batch with every event on the tiled engine (bucket table with one tiled row):
    tile side R (tq 0 or 4), 1.25 R, 1.5 R and 4 R, with each kernel variant (tile
    tables of 512, 384, 320 and 256 tiles): the history depends on neither."""
    import fastjet_gpu_amd as fj
    saved = engine.buckets()
    try:
        engine.set_buckets([(1000000000, 2, 64, tq, var)])
        assert engine.buckets() == [(1000000000, 2, 64, tq, var)]
        p4, off = fj.gen_workload(3, 64, first=100 + tq)
        ne = len(off) - 1
        ps = np.array([(-1, 0, 1)[e % 3] for e in range(ne)], np.int32)
        Rs = np.array([(0.4, 0.8, 1.0, 0.6)[e % 4] for e in range(ne)], np.float64)
        print(f"[gpu parity] tiled batch workload3 x{ne} tile side {max(tq, 4) / 4} R variant {var}", flush=True)
        h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, ps, Rs)
        for e in range(ne):
            a, b = off[e], off[e + 1]
            want = orc.n2plain(p4[a:b], int(ps[e]), float(Rs[e]))
            got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
            assert same(want, got), f"tiled batch event {e} (N={b - a}): {first_diff(want, got)}"
        a, b = off[0], off[1]
        alone = engine.cluster_event(p4[a:b], int(ps[0]), float(Rs[0]))  # AUTO: the tiled row, same tile side
        assert same(alone, (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[0]]))
    finally:
        engine.set_buckets(saved)


@pytest.mark.gpu
def test_bucket_table_validation(engine):
    import fastjet_gpu_amd as fj
    saved = engine.buckets()
    bad = [[(1000000000, 2, 64, 3, 0)], [(1000000000, 2, 64, 17, 0)], [(1000000000, 2, 128, 0, 0)],
           [(1000000000, 2, 64, 0, 4)], [(1000000000, 2, 64, 0, -1)],
           [(300, 1, 64, 4, 0)], [(256, 1, 64, 4, 0), (128, 2, 64, 0, 0)], [(64, 1, 64, 3, 0)], []]
    for rows in bad:
        with pytest.raises(fj.FjmiError):
            engine.set_buckets(rows)
        assert engine.buckets() == saved


@pytest.mark.gpu
def test_forced_dense_beyond_table(engine):
    """engine=dense for an event larger than every dense table row runs the smallest
    compiled dense kernel that holds it."""
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(4, 1)[:3000]
    check_event(engine, ev, -1, 0.4, "forced dense N=3000", 1)
    with pytest.raises(fj.FjmiError):
        engine.cluster_event(fj.gen_event(4, 2)[:9000], -1, 0.4, engine=1)


@pytest.mark.gpu
def test_batch_matches_oracle(engine):
    """Mixed multiplicities, one graph launch, per-event algorithm and radius."""
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 96)
    ne = len(off) - 1
    ps = np.array([(-1, 0, 1)[e % 3] for e in range(ne)], np.int32)
    Rs = np.array([(0.4, 0.8, 1.0, 0.6)[e % 4] for e in range(ne)], np.float64)
    print(f"[gpu parity] batch workload3 x{ne} events, total {len(p4)} particles", flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, ps, Rs)
    for e in range(ne):
        a, b = off[e], off[e + 1]
        want = orc.n2plain(p4[a:b], int(ps[e]), float(Rs[e]))
        got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
        assert same(want, got), f"batch event {e} (N={b - a}): {first_diff(want, got)}"


@pytest.mark.gpu
@pytest.mark.parametrize("var", [1, 0, 3])
def test_stretched_tiles(engine, var):
    """Events whose tile grid exceeds the kernel's tile table, so that tiles get
    taller than R in rapidity: R = 0.05 and 0.1 (125 and 62 columns in phi) and
    rapidities up to +-9 at R = 0.4 (45 rows at side R), with the 384-, 512- and
    256-tile variants (the last leaves 2 rows at R = 0.05)."""
    import fastjet_gpu_amd as fj
    saved = engine.buckets()
    try:
        engine.set_buckets([(1000000000, 2, 64, 0, var)])
        p4, off = fj.gen_workload(3, 6, first=4200)
        pr = [(-1, 0.05), (1, 0.1), (0, 0.1), (-1, 0.1), (1, 0.05), (0, 0.05)]
        cases = [(p4[off[e]:off[e + 1]], p, R) for e, (p, R) in enumerate(pr)]
        rng = np.random.default_rng(5)
        for n in (300, 900):
            pt = 0.3 + rng.exponential(1.0, n)
            eta = rng.uniform(-9.0, 9.0, n)
            phi = rng.uniform(0.0, 2.0 * np.pi, n)
            wide = np.stack([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)], 1)
            cases += [(wide, -1, 0.4), (wide, 1, 0.4)]
        for k, (ev, p, R) in enumerate(cases):
            check_event(engine, np.ascontiguousarray(ev), p, R, f"stretched tiles, variant {var}, case {k}", eng=2)
            check_event(engine, np.ascontiguousarray(ev), p, R, f"stretched tiles, variant {var}, case {k} (auto)")
    finally:
        engine.set_buckets(saved)


@pytest.mark.gpu
@pytest.mark.parametrize("workload,ne,first", [(3, 512, 3000), (5, 128, 500), (2, 64, 10)])
def test_batch_mixed_algorithms_and_radii(engine, workload, ne, first):
    """One batch whose events each draw their algorithm from {kt, C/A, anti-kt} and R from
    {0.2, 0.4, 0.7, 1.0, 1.5} (seeded): every event bit-exact against the oracle."""
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(workload, ne, first=first)
    rng = np.random.default_rng(11 + workload)
    ps = rng.choice([-1, 0, 1], size=ne).astype(np.int32)
    Rs = rng.choice([0.2, 0.4, 0.7, 1.0, 1.5], size=ne).astype(np.float64)
    print(f"[gpu parity] workload {workload} events {first}..{first + ne - 1}, {len(p4)} particles, mixed p and R",
          flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, ps, Rs)
    for e in range(ne):
        a, b = off[e], off[e + 1]
        pe, re = int(ps[e]), float(Rs[e])
        want = orc.n2plain(p4[a:b], pe, re) if b - a <= 3000 else orc.tiled(p4[a:b], pe, re)
        got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
        assert same(want, got), f"event {first + e} (N={b - a}, p={pe}, R={re}): {first_diff(want, got)}"


@pytest.mark.gpu
def test_batch_alone_equals_in_mix(engine):
    """Each event's result is the same clustered alone and inside a batch."""
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 24, first=500)
    print(f"[gpu parity] batch vs alone, {len(off) - 1} events", flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, -1, 0.4)
    for e in range(len(off) - 1):
        a, b = off[e], off[e + 1]
        alone = engine.cluster_event(p4[a:b], -1, 0.4)
        assert same(alone, (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])), f"event {e}"


@pytest.mark.gpu
def test_pseudojet_dropin_layout(engine):
    """cluster.h contract: in place, isJet slots hold the jets, jet promoted with n live sits in slot n-1."""
    import torch
    ev = orc.read_events(os.path.join(GOLDEN, "808_cms.dat"))[0]
    n = len(ev)
    rec = np.zeros(n, dtype=[("index", "<i4"), ("isJet", "u1"), ("pad", "u1", 3), ("p4", "<f8", 4)])
    assert rec.itemsize == 40
    rec["index"] = np.arange(n)
    rec["p4"] = ev
    raw = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    print(f"[gpu parity] drop-in PseudoJet N={n}", flush=True)
    engine.cluster_pseudojets(raw.data_ptr(), n, -1, 0.4)
    torch.cuda.synchronize()
    out = raw.cpu().numpy().view(rec.dtype)
    h1, h2, hd, jets = orc.n2plain(ev, -1, 0.4)
    assert np.array_equal(out["index"], np.arange(n))
    jet_slots = [n - 1 - s for s in range(n) if h2[s] == -1]  # beam step s had n-s live jets
    assert sorted(np.nonzero(out["isJet"])[0].tolist()) == sorted(jet_slots)
    for j, slot in enumerate(jet_slots):
        assert out["p4"][slot].tobytes() == jets[j].tobytes()


@pytest.mark.gpu
def test_one_event_calls_on_several_streams(engine):
    """One-event device calls enqueued on three streams back to back, with no
    synchronisation between them: each stream has its own scratch and status
    word, so the overlapping launches do not share state; every result is
    bit-exact and each stream's check reports no failure."""
    import torch
    import fastjet_gpu_amd as fj
    evs = [fj.gen_event(3, 7), fj.gen_event(3, 11), fj.gen_event(2, 0), fj.gen_event(3, 12)]
    streams = [torch.cuda.Stream(), torch.cuda.Stream(), torch.cuda.Stream()]
    bufs = []
    for ev in evs:
        n = len(ev)
        bufs.append(dict(p4=torch.from_numpy(np.ascontiguousarray(ev)).cuda(),
                         h1=torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                         h2=torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                         hd=torch.zeros(n, dtype=torch.float64, device="cuda"),
                         jets=torch.zeros(4 * n, dtype=torch.float64, device="cuda"),
                         nj=torch.full((1,), -1, dtype=torch.int32, device="cuda")))
    torch.cuda.synchronize()
    print(f"[gpu parity] {len(evs)} one-event calls on {len(streams)} streams, N={[len(e) for e in evs]}", flush=True)
    for k, (ev, b) in enumerate(zip(evs, bufs)):
        st = streams[k % len(streams)]
        engine.cluster_event_device(b["p4"].data_ptr(), len(ev), -1, 0.4, b["h1"].data_ptr(), b["h2"].data_ptr(),
                                    b["hd"].data_ptr(), b["jets"].data_ptr(), b["nj"].data_ptr(), st.cuda_stream)
    for st in streams:
        engine.stream_check(st.cuda_stream)
    for ev, b in zip(evs, bufs):
        nj = int(b["nj"].cpu()[0])
        got = (b["h1"].cpu().numpy(), b["h2"].cpu().numpy(), b["hd"].cpu().numpy(),
               b["jets"].cpu().numpy().reshape(-1, 4)[:nj])
        want = orc.n2plain(ev, -1, 0.4)
        assert same(want, got), f"N={len(ev)}: {first_diff(want, got)}"


@pytest.mark.gpu
def test_stream_check_without_failures(engine):
    """fjmi_stream_check on a stream that never ran a one-event call, and after a
    successful one, returns without error."""
    import torch
    s = torch.cuda.Stream()
    engine.stream_check(s.cuda_stream)
    ev = orc.read_events(os.path.join(GOLDEN, "354_fastjet.dat"))[0]
    n = len(ev)
    d = torch.from_numpy(ev).cuda()
    out = [torch.zeros(n, dtype=torch.int32, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda"),
           torch.zeros(n, dtype=torch.float64, device="cuda"), torch.zeros(4 * n, dtype=torch.float64, device="cuda"),
           torch.zeros(1, dtype=torch.int32, device="cuda")]
    torch.cuda.synchronize()
    engine.cluster_event_device(d.data_ptr(), n, 1, 0.4, *[t.data_ptr() for t in out], s.cuda_stream)
    engine.stream_check(s.cuda_stream)
    engine.stream_check(0)
