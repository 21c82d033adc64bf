"""GPU parity of the cooperative tiled kernel (tiled_coop_kernel: 8 waves per
event share the member passes and orphan scans of gathers of several chunks),
which the default table's top row (N > 4096) runs while a bucket holds at most
one event per CU, and one-event calls run for N R^2 > 1200.  Bar: bit-exact
history and jets against the oracle (its tiled restatement above N = 6000) and
against the one-wave kernel."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_parity import first_diff, same


def _batch(engine, p4, off, p, R):
    import torch
    tot, ne = len(p4), len(off) - 1
    d = dict(p4=torch.from_numpy(np.ascontiguousarray(p4)).cuda(), off=torch.from_numpy(off).cuda(),
             h1=torch.full((tot,), -7, dtype=torch.int32, device="cuda"),
             h2=torch.full((tot,), -7, dtype=torch.int32, device="cuda"),
             hd=torch.zeros(tot, dtype=torch.float64, device="cuda"),
             j=torch.zeros(4 * tot, dtype=torch.float64, device="cuda"),
             nj=torch.full((ne,), -7, dtype=torch.int32, device="cuda"))
    torch.cuda.synchronize()
    b = engine.batch_create(off, d["p4"].data_ptr(), d["off"].data_ptr(), 0, 0, p, R, d["h1"].data_ptr(),
                            d["h2"].data_ptr(), d["hd"].data_ptr(), d["j"].data_ptr(), d["nj"].data_ptr())
    try:
        b.launch()
        b.sync()
        stats = b.stats()
    finally:
        b.close()
    out = (d["h1"].cpu().numpy(), d["h2"].cpu().numpy(), d["hd"].cpu().numpy(),
           d["j"].cpu().numpy().reshape(-1, 4), d["nj"].cpu().numpy())
    return out, stats


@pytest.mark.gpu
@pytest.mark.parametrize("p", [-1, 1])
def test_high_pileup_batch_cooperative(engine, p):
    """The high-pileup batch (64 events, N ~ U[5000, 20000], R = 0.4) on the
    default table: its top row runs 8 waves per event; equal to one wave per
    event bit for bit, and to the oracle on three events."""
    import fastjet_gpu_amd as fj
    saved = engine.buckets()
    p4, off = fj.gen_workload(4, 64)
    print(f"[gpu coop] high-pileup batch x64 p={p}: default table", flush=True)
    got, stats = _batch(engine, p4, off, p, 0.4)
    assert any(s["threads"] == 512 for s in stats), stats
    try:
        engine.set_buckets([(4096, 2, 64, 4, 3), (1000000000, 2, 64, 4, 1)])
        print(f"[gpu coop] high-pileup batch x64 p={p}: one wave per event", flush=True)
        ref, stats1 = _batch(engine, p4, off, p, 0.4)
        assert all(s["threads"] == 64 for s in stats1)
    finally:
        engine.set_buckets(saved)
    for x, y in zip(ref, got):
        assert x.tobytes() == y.tobytes()
    h1, h2, hd, jets, nj = got
    for e in (0, 17, 40):
        a, b = off[e], off[e + 1]
        print(f"[gpu coop] oracle check event {e} N={b - a}", flush=True)
        want = orc.tiled(p4[a:b], p, 0.4)
        res = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
        assert same(want, res), f"event {e}: {first_diff(want, res)}"


@pytest.mark.gpu
@pytest.mark.parametrize("R,p", [(0.8, -1), (1.0, 0), (0.4, 1)])
def test_single_large_events(engine, R, p):
    """One-event calls with N R^2 > 1200 take the cooperative kernel (latency rule)."""
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(4, 11)
    if len(ev) * R * R <= 1200:
        ev = fj.gen_event(4, 0)
    print(f"[gpu coop] one event N={len(ev)} R={R} p={p}", flush=True)
    got = engine.cluster_event(ev, p, R)
    want = orc.tiled(ev, p, R)
    assert same(want, got), first_diff(want, got)


@pytest.mark.gpu
def test_many_large_events_fall_back_to_one_wave(engine):
    """A bucket with more events than CUs runs one wave per event (more events
    in flight) whatever its row says; results unchanged."""
    import fastjet_gpu_amd as fj
    saved = engine.buckets()
    try:
        engine.set_buckets([(300, 2, 64, 4, 3), (1000000000, 2, 512, 4, 1)])
        p4, off = fj.gen_workload(3, 600, first=3000)
        sizes = np.diff(off)
        print(f"[gpu coop] {int((sizes > 300).sum())} events above N = 300 on a 512-thread row", flush=True)
        got, stats = _batch(engine, p4, off, -1, 1.0)
        top = [s for s in stats if s["max_n"] == 1000000000][0]
        assert top["threads"] == 64 and top["n_events"] > 256
        engine.set_buckets([(300, 2, 64, 4, 3), (1000000000, 2, 512, 4, 1)])
        p4s, offs = p4[:off[40]], off[:41]
        print("[gpu coop] the same row with 40 events: 8 waves per event", flush=True)
        got2, stats2 = _batch(engine, p4s, offs, -1, 1.0)
        assert [s for s in stats2 if s["max_n"] == 1000000000][0]["threads"] == 512
    finally:
        engine.set_buckets(saved)
    h1, h2, hd, jets, nj = got
    g1, g2, gd, gj, gn = got2
    for e in range(40):
        a, b = off[e], off[e + 1]
        assert (h1[a:b].tobytes() == g1[a:b].tobytes() and hd[a:b].tobytes() == gd[a:b].tobytes()
                and nj[e] == gn[e] and jets[a:a + nj[e]].tobytes() == gj[a:a + gn[e]].tobytes()), f"event {e}"
    for e in (0, 7, 123, 599):
        a, b = off[e], off[e + 1]
        want = orc.n2plain(p4[a:b], -1, 1.0)
        assert same(want, (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])), f"event {e}"
