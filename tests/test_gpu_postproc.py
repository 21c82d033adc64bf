"""GPU parity of the results derived on the device (postproc.hip, through the C
ABI) against the oracle's restatements, bit for bit: inclusive jets above ptmin
sorted by pT (main.cc:309-327), on packed jets of a batch and on the drop-in
PseudoJet records; FastJet's history view (child, max_dij_so_far); FastJet's
exclusive jets by count and by dcut."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import GOLDEN


def _batch(engine):
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 48, first=700)
    ne = len(off) - 1
    ps = np.array([(-1, 0, 1)[e % 3] for e in range(ne)], np.int32)
    Rs = np.array([(0.4, 0.7, 1.0)[e % 3] for e in range(ne)], np.float64)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, ps, Rs)
    return p4, off, ps, Rs, h1, h2, hd, jets, nj


def _t(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.gpu
def test_inclusive_jets_batch(engine):
    import torch
    p4, off, ps, Rs, h1, h2, hd, jets, nj = _batch(engine)
    ne = len(off) - 1
    dj, dn, doff = _t(torch, jets), _t(torch, nj.astype(np.int32)), _t(torch, off)
    for ptmin in (0.0, 1.0, 5.0):
        out = torch.zeros_like(dj)
        on = torch.full((ne,), -2, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        print(f"[gpu postproc] inclusive jets, {ne} events, ptmin {ptmin}", flush=True)
        engine.inclusive_jets(ne, doff.data_ptr(), dj.data_ptr(), dn.data_ptr(), ptmin, out.data_ptr(), on.data_ptr())
        torch.cuda.synchronize()
        o, n = out.cpu().numpy().reshape(-1, 4), on.cpu().numpy()
        for e in range(ne):
            a = off[e]
            want = orc.inclusive_jets(jets[a:a + nj[e]], ptmin)
            assert n[e] == len(want)
            assert o[a:a + n[e]].tobytes() == want.tobytes(), f"event {e}"


@pytest.mark.gpu
def test_history_view_batch(engine):
    import torch
    p4, off, ps, Rs, h1, h2, hd, jets, nj = _batch(engine)
    ne, tot = len(off) - 1, len(p4)
    child = torch.zeros(2 * tot, dtype=torch.int32, device="cuda")
    md = torch.zeros(tot, dtype=torch.float64, device="cuda")
    d1, d2, dd, doff = _t(torch, h1), _t(torch, h2), _t(torch, hd), _t(torch, off)
    torch.cuda.synchronize()
    print(f"[gpu postproc] history view, {ne} events", flush=True)
    engine.history_view(ne, doff.data_ptr(), 0, d1.data_ptr(), d2.data_ptr(), dd.data_ptr(), child.data_ptr(),
                        md.data_ptr())
    torch.cuda.synchronize()
    c, m = child.cpu().numpy(), md.cpu().numpy()
    for e in range(ne):
        a, b = off[e], off[e + 1]
        wc, wm = orc.history_children(h1[a:b], h2[a:b], hd[a:b])
        assert np.array_equal(c[2 * a:2 * b], wc) and m[a:b].tobytes() == wm.tobytes(), f"event {e}"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [0, 1])
def test_exclusive_jets_batch(engine, mode):
    import torch
    p4, off, ps, Rs, h1, h2, hd, jets, nj = _batch(engine)
    ne, tot = len(off) - 1, len(p4)
    dp4, d1, d2, dd, doff = _t(torch, p4), _t(torch, h1), _t(torch, h2), _t(torch, hd), _t(torch, off)
    work = torch.zeros(8 * tot, dtype=torch.float64, device="cuda")
    params = (0, 1, 2, 5, 60, 3000) if mode == 0 else (0.0, 1e-3, 0.1, 1.0, 25.0, 1e9)
    for param in params:
        out = torch.zeros(4 * tot, dtype=torch.float64, device="cuda")
        on = torch.full((ne,), -2, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        print(f"[gpu postproc] exclusive jets mode {mode} param {param}", flush=True)
        engine.exclusive_jets(ne, doff.data_ptr(), 0, dp4.data_ptr(), d1.data_ptr(), d2.data_ptr(), dd.data_ptr(),
                              mode, param, work.data_ptr(), out.data_ptr(), on.data_ptr())
        torch.cuda.synchronize()
        o, n = out.cpu().numpy().reshape(-1, 4), on.cpu().numpy()
        for e in range(ne):
            a, b = off[e], off[e + 1]
            if mode == 0 and param > b - a:
                assert n[e] == -1
                continue
            want = orc.exclusive_jets(p4[a:b], h1[a:b], h2[a:b], hd[a:b], mode, param)
            assert n[e] == len(want) and o[a:a + n[e]].tobytes() == want.tobytes(), f"event {e} param {param}"


@pytest.mark.gpu
def test_single_event_forms(engine):
    """n_events = 1 without offsets, on the 808-particle reference event."""
    import torch
    ev = orc.read_events(os.path.join(GOLDEN, "808_cms.dat"))[0]
    n = len(ev)
    h1, h2, hd, jets = engine.cluster_event(ev, 1, 0.4)
    d1, d2, dd, dp4 = _t(torch, h1), _t(torch, h2), _t(torch, hd), _t(torch, ev)
    child = torch.zeros(2 * n, dtype=torch.int32, device="cuda")
    md = torch.zeros(n, dtype=torch.float64, device="cuda")
    work = torch.zeros(8 * n, dtype=torch.float64, device="cuda")
    out = torch.zeros(4 * n, dtype=torch.float64, device="cuda")
    on = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    engine.history_view(1, 0, n, d1.data_ptr(), d2.data_ptr(), dd.data_ptr(), child.data_ptr(), md.data_ptr())
    engine.exclusive_jets(1, 0, n, dp4.data_ptr(), d1.data_ptr(), d2.data_ptr(), dd.data_ptr(), 0, 4,
                          work.data_ptr(), out.data_ptr(), on.data_ptr())
    torch.cuda.synchronize()
    wc, wm = orc.history_children(h1, h2, hd)
    assert np.array_equal(child.cpu().numpy(), wc) and md.cpu().numpy().tobytes() == wm.tobytes()
    want = orc.exclusive_jets(ev, h1, h2, hd, 0, 4)
    assert int(on.cpu()[0]) == 4 and out.cpu().numpy().reshape(-1, 4)[:4].tobytes() == want.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("ptmin", [0.0, 1.0])
def test_inclusive_pseudojets(engine, ptmin):
    """After the drop-in clustering of PseudoJet records: the isJet records above
    ptmin, whole and sorted, as the CLI now reads them back."""
    import torch
    ev = orc.read_events(os.path.join(GOLDEN, "808_cms.dat"))[0]
    n = len(ev)
    dt = np.dtype([("index", "<i4"), ("isJet", "u1"), ("pad", "u1", 3), ("p4", "<f8", 4)])
    rec = np.zeros(n, dtype=dt)
    rec["index"] = np.arange(n)
    rec["p4"] = ev
    raw = _t(torch, rec.view(np.uint8))
    out = torch.zeros_like(raw)
    cnt = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    engine.cluster_pseudojets(raw.data_ptr(), n, -1, 0.4)
    engine.inclusive_pseudojets(raw.data_ptr(), n, ptmin, out.data_ptr(), cnt.data_ptr())
    torch.cuda.synchronize()
    k = int(cnt.cpu()[0])
    got = out.cpu().numpy().view(dt)[:k]
    want = orc.inclusive_jets(orc.n2plain(ev, -1, 0.4)[3], ptmin)
    assert k == len(want) and got["p4"].tobytes() == want.tobytes()
    assert np.all(got["isJet"] == 1)
