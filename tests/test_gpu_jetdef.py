"""GPU parity of general jet definitions: any real genkt exponent p and FastJet's
recombination schemes, on the tiled engine's tiled_gen_kernel variants, through
the C ABI (fjmi_cluster_event_def, fjmi_cluster_event_device_def,
fjmi_batch_create_def).

Bar: bit-exact merge history (parents, dij as raw bits) and bit-exact jets against
the oracle's restatement of the same definition (oracle/fj_oracle.c, *_def).  The
reference has p in {1, 0, -1} with the E-scheme only (cluster.h:4,
tri_matrix.cu:247-250); for those definitions the general kernels must also
agree bit for bit with the standard ones.  Every test prints its case before
touching the GPU.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import GOLDEN
from tests.test_gpu_parity import first_diff, same

SCHEMES = list(range(9))  # E, pt, pt2, Et, Et2, BIpt, BIpt2, WTA_pt, WTA_modp


def check_def(engine, p4, p, scheme, R, label):
    print(f"[gpu jetdef] {label} N={len(p4)} p={p} scheme={scheme} R={R}", flush=True)
    want = orc.cluster_def(p4, p, scheme, R, "n2plain" if len(p4) <= 6000 else "tiled")
    got = engine.cluster_event_def(p4, p, scheme, R)
    assert same(want, got), f"{label}: first differing step {first_diff(want, got)}"


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 0, -1])
@pytest.mark.parametrize("R", [0.4, 1.0])
def test_standard_definitions_equal_standard_kernels(engine, p, R):
    """p in {1, 0, -1} with the E-scheme: the general kernel gives the standard kernel's bits."""
    ev = orc.read_events(os.path.join(GOLDEN, "808_cms.dat"))[0]
    print(f"[gpu jetdef] 808_cms.dat p={p} E-scheme R={R}: general vs standard kernel", flush=True)
    got = engine.cluster_event_def(ev, float(p), 0, R)
    std = engine.cluster_event(ev, p, R, engine=2)
    assert same(std, got), first_diff(std, got)
    assert same(orc.n2plain(ev, p, R), got)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("p", [-1.0, 0.0, 1.0, 0.5, -0.75, 2.0])
def test_schemes_reference_file(engine, scheme, p):
    ev = orc.read_events(os.path.join(GOLDEN, "354_fastjet.dat"))[0]
    check_def(engine, ev, p, scheme, 0.6, "354_fastjet.dat")


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", SCHEMES)
def test_schemes_generated(engine, scheme):
    """Generated events of the 4096-event workload, one exponent per scheme (the
    kernel's larger-gather paths at R = 1.0 and 1.5)."""
    import fastjet_gpu_amd as fj
    ps = [0.3, -1.7, 1.0, -0.2, 0.8, -1.0, 1.4, -3.0, 0.0]
    for k, R in enumerate((0.4, 1.0, 1.5)):
        ev = fj.gen_event(3, 40 + 3 * scheme + k)
        check_def(engine, ev, ps[scheme], scheme, R, f"workload3 event {40 + 3 * scheme + k}")


@pytest.mark.gpu
@pytest.mark.parametrize("p,scheme", [(0.5, 1), (-1.0, 7), (1.0, 2)])
def test_high_multiplicity(engine, p, scheme):
    """A high-pileup event (several-chunk gathers), checked against the oracle's tiled restatement."""
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(4, 3)
    check_def(engine, ev, p, scheme, 0.4, "workload4 event 3")


@pytest.mark.gpu
def test_edge_events(engine):
    """Single particles, pairs, duplicates and an exact-tie lattice under every scheme."""
    lat = []
    for i in range(10):
        for j in range(10):
            eta, phi, pt = -1.0 + 0.25 * i, 0.25 * j, 1.0 + 0.125 * ((i * 7 + j * 3) % 5)
            lat.append([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)])
    events = {
        "single": np.array([[3.0, 1.0, 0.5, 4.0]]),
        "pair": np.array([[3.0, 1.0, 0.5, 4.0], [2.9, 1.1, 0.4, 3.8]]),
        "duplicates": np.array([[1.0, 2.0, 0.5, 3.0]] * 5 + [[-1.0, 0.2, 2.0, 3.0]] * 3),
        "lattice": np.array(lat),
    }
    for name, ev in events.items():
        for scheme in SCHEMES:
            for p in (-1.0, 0.0, 1.0, 0.4):
                check_def(engine, ev, p, scheme, 0.7, name)
    h1, h2, hd, jets = engine.cluster_event_def(np.zeros((0, 4)), 0.5, 1, 0.4)
    assert len(h1) == 0 and len(jets) == 0


@pytest.mark.gpu
def test_invalid_definitions_refused(engine):
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(1, 0)
    for p, scheme in ((float("nan"), 0), (float("inf"), 0), (1e301, 1), (0.5, -1), (0.5, 9)):
        print(f"[gpu jetdef] refused definition p={p} scheme={scheme}", flush=True)
        with pytest.raises(fj.FjmiError):
            engine.cluster_event_def(ev, p, scheme, 0.4)


def _device_batch(engine, p4, off, defs, Rs):
    import torch
    import fastjet_gpu_amd as fj
    tot, ne = len(p4), len(off) - 1
    d = dict(p4=torch.from_numpy(np.ascontiguousarray(p4)).cuda(), off=torch.from_numpy(off).cuda(),
             R=torch.from_numpy(Rs).cuda(), h1=torch.full((tot,), -7, dtype=torch.int32, device="cuda"),
             h2=torch.full((tot,), -7, dtype=torch.int32, device="cuda"),
             hd=torch.zeros(tot, dtype=torch.float64, device="cuda"),
             jets=torch.zeros(4 * tot, dtype=torch.float64, device="cuda"),
             nj=torch.full((ne,), -7, dtype=torch.int32, device="cuda"))
    if defs is not None:
        rec = np.zeros(ne, fj.JETDEF_DTYPE)
        rec["p"], rec["scheme"] = defs
        d["defs"] = torch.from_numpy(rec.view(np.uint8).copy()).cuda()
    torch.cuda.synchronize()
    return d


@pytest.mark.gpu
def test_batch_per_event_definitions(engine):
    """One graph launch with a definition and a radius per event, twice (plan reuse)."""
    import torch
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 72, first=500)
    ne = len(off) - 1
    rng = np.random.default_rng(7)
    ps = np.round(rng.uniform(-2.0, 2.0, ne), 3)
    ps[:3] = (-1.0, 0.0, 1.0)
    schemes = np.array([e % 9 for e in range(ne)], np.int32)
    Rs = np.array([(0.4, 0.8, 1.0, 0.6)[e % 4] for e in range(ne)], np.float64)
    print(f"[gpu jetdef] batch workload3 x{ne} per-event definitions", flush=True)
    d = _device_batch(engine, p4, off, (ps, schemes), Rs)
    b = engine.batch_create_def(off, d["p4"].data_ptr(), d["off"].data_ptr(), d["defs"].data_ptr(), 0.0, 0,
                                d["R"].data_ptr(), 0.0, d["h1"].data_ptr(), d["h2"].data_ptr(), d["hd"].data_ptr(),
                                d["jets"].data_ptr(), d["nj"].data_ptr())
    try:
        for rep in range(2):
            b.launch()
            b.sync()
            h1, h2, hd = d["h1"].cpu().numpy(), d["h2"].cpu().numpy(), d["hd"].cpu().numpy()
            jets, nj = d["jets"].cpu().numpy().reshape(-1, 4), d["nj"].cpu().numpy()
            for e in range(ne):
                a, c = off[e], off[e + 1]
                want = orc.cluster_def(p4[a:c], ps[e], int(schemes[e]), Rs[e])
                got = (h1[a:c], h2[a:c], hd[a:c], jets[a:a + nj[e]])
                assert same(want, got), f"launch {rep} event {e} p={ps[e]} scheme={schemes[e]}: {first_diff(want, got)}"
    finally:
        b.close()
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_batch_uniform_definition_and_invalid_event(engine):
    """One definition for the batch; then a per-event table with one invalid
    definition: that event alone fails (njets -1) and the sync reports it."""
    import torch
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 24, first=900)
    ne = len(off) - 1
    Rs = np.full(ne, 0.5)
    print(f"[gpu jetdef] batch workload3 x{ne} uniform p=0.25 scheme=pt2", flush=True)
    d = _device_batch(engine, p4, off, None, Rs)
    b = engine.batch_create_def(off, d["p4"].data_ptr(), d["off"].data_ptr(), 0, 0.25, fj.SCHEME_PT2, 0, 0.5,
                                d["h1"].data_ptr(), d["h2"].data_ptr(), d["hd"].data_ptr(), d["jets"].data_ptr(),
                                d["nj"].data_ptr())
    try:
        b.launch()
        b.sync()
    finally:
        b.close()
    h1, h2, hd = d["h1"].cpu().numpy(), d["h2"].cpu().numpy(), d["hd"].cpu().numpy()
    jets, nj = d["jets"].cpu().numpy().reshape(-1, 4), d["nj"].cpu().numpy()
    for e in range(ne):
        a, c = off[e], off[e + 1]
        want = orc.cluster_def(p4[a:c], 0.25, fj.SCHEME_PT2, 0.5)
        assert same(want, (h1[a:c], h2[a:c], hd[a:c], jets[a:a + nj[e]])), f"event {e}"

    print(f"[gpu jetdef] batch x{ne} with one invalid per-event scheme", flush=True)
    schemes = np.zeros(ne, np.int32)
    schemes[5] = 42
    d = _device_batch(engine, p4, off, (np.full(ne, -1.0), schemes), Rs)
    b = engine.batch_create_def(off, d["p4"].data_ptr(), d["off"].data_ptr(), d["defs"].data_ptr(), 0.0, 0,
                                d["R"].data_ptr(), 0.0, d["h1"].data_ptr(), d["h2"].data_ptr(), d["hd"].data_ptr(),
                                d["jets"].data_ptr(), d["nj"].data_ptr())
    try:
        b.launch()
        with pytest.raises(fj.FjmiError):
            b.sync()
    finally:
        b.close()
    nj = d["nj"].cpu().numpy()
    assert nj[5] == -1
    h1, h2, hd = d["h1"].cpu().numpy(), d["h2"].cpu().numpy(), d["hd"].cpu().numpy()
    jets = d["jets"].cpu().numpy().reshape(-1, 4)
    for e in (0, 4, 6, ne - 1):
        a, c = off[e], off[e + 1]
        assert same(orc.n2plain(p4[a:c], -1, 0.5), (h1[a:c], h2[a:c], hd[a:c], jets[a:a + nj[e]])), f"event {e}"


@pytest.mark.gpu
def test_device_entry_point(engine):
    """fjmi_cluster_event_device_def on a stream, then fjmi_stream_check."""
    import torch
    import fastjet_gpu_amd as fj
    ev = fj.gen_event(2, 0)
    n = len(ev)
    print(f"[gpu jetdef] device entry N={n} p=-0.5 scheme=WTA_pt R=0.8", flush=True)
    s = torch.cuda.Stream()
    dp = torch.from_numpy(np.ascontiguousarray(ev)).cuda()
    out = [torch.zeros(n, dtype=torch.int32, device="cuda"), torch.zeros(n, dtype=torch.int32, device="cuda"),
           torch.zeros(n, dtype=torch.float64, device="cuda"), torch.zeros(4 * n, dtype=torch.float64, device="cuda"),
           torch.zeros(1, dtype=torch.int32, device="cuda")]
    torch.cuda.synchronize()
    engine.cluster_event_device_def(dp.data_ptr(), n, -0.5, fj.SCHEME_WTA_PT, 0.8, *[t.data_ptr() for t in out],
                                    stream=s.cuda_stream)
    engine.stream_check(s.cuda_stream)
    nj = int(out[4].cpu().numpy()[0])
    got = (out[0].cpu().numpy(), out[1].cpu().numpy(), out[2].cpu().numpy(),
           out[3].cpu().numpy().reshape(-1, 4)[:nj])
    want = orc.cluster_def(ev, -0.5, fj.SCHEME_WTA_PT, 0.8)
    assert same(want, got), first_diff(want, got)


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [0, 1, 4, 7, 8])
def test_history_momenta(engine, scheme):
    """The momentum each merge step stores (FastJet's ClusterSequence jets), bit for bit."""
    ev = orc.read_events(os.path.join(GOLDEN, "354_fastjet.dat"))[0]
    print(f"[gpu jetdef] history momenta 354_fastjet.dat p=0.5 scheme={scheme} R=0.8", flush=True)
    got = engine.cluster_event_def(ev, 0.5, scheme, 0.8, hist_jets=True)
    want = orc.cluster_def_hist(ev, 0.5, scheme, 0.8)
    assert same(want[:4], got[:4]), first_diff(want, got)
    merge = want[1] >= 0
    assert want[4][merge].tobytes() == got[4][merge].tobytes()


@pytest.mark.gpu
def test_exclusive_jets_under_schemes(engine):
    """A batch with per-event definitions writes each merge step's momentum;
    fjmi_exclusive_jets_def then gives the oracle's exclusive jets by count and by dcut."""
    import torch
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 18, first=1200)
    ne, tot = len(off) - 1, len(p4)
    ps = np.array([(1.0, 0.5, -1.0, 0.0, 2.0, -0.5)[e % 6] for e in range(ne)])
    schemes = np.array([(0, 1, 3, 5, 7, 8, 2, 4, 6)[e % 9] for e in range(ne)], np.int32)
    Rs = np.full(ne, 1.0)
    print(f"[gpu jetdef] exclusive jets under schemes: batch x{ne}", flush=True)
    d = _device_batch(engine, p4, off, (ps, schemes), Rs)
    hj = torch.zeros(4 * tot, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    b = engine.batch_create_def(off, d["p4"].data_ptr(), d["off"].data_ptr(), d["defs"].data_ptr(), 0.0, 0,
                                d["R"].data_ptr(), 0.0, d["h1"].data_ptr(), d["h2"].data_ptr(), d["hd"].data_ptr(),
                                d["jets"].data_ptr(), d["nj"].data_ptr(), hj.data_ptr())
    try:
        b.launch()
        b.sync()
    finally:
        b.close()
    h1, h2, hd = d["h1"].cpu().numpy(), d["h2"].cpu().numpy(), d["hd"].cpu().numpy()
    hjn = hj.cpu().numpy().reshape(-1, 4)
    ref = []
    for e in range(ne):
        a, c = off[e], off[e + 1]
        want = orc.cluster_def_hist(p4[a:c], ps[e], int(schemes[e]), 1.0)
        assert (np.array_equal(want[0], h1[a:c]) and np.array_equal(want[1], h2[a:c])
                and want[2].tobytes() == hd[a:c].tobytes()), f"event {e}"
        merge = want[1] >= 0
        assert want[4][merge].tobytes() == hjn[a:c][merge].tobytes(), f"event {e} history momenta"
        ref.append(want)
    off_d = d["off"]
    for mode, params in ((0, (0, 1, 4, 30)), (1, (0.0, 0.05, 1.0, 1e9))):
        for param in params:
            for sch in range(9):  # each call takes one scheme: check the events clustered under it
                evs = [e for e in range(ne) if schemes[e] == sch]
                if not evs:
                    continue
                out = torch.zeros(4 * tot, dtype=torch.float64, device="cuda")
                on = torch.full((ne,), -2, dtype=torch.int32, device="cuda")
                torch.cuda.synchronize()
                engine.exclusive_jets_def(ne, off_d.data_ptr(), 0, d["p4"].data_ptr(), d["h1"].data_ptr(),
                                          d["h2"].data_ptr(), d["hd"].data_ptr(), hj.data_ptr(), sch, mode, param,
                                          out.data_ptr(), on.data_ptr())
                torch.cuda.synchronize()
                o, n = out.cpu().numpy().reshape(-1, 4), on.cpu().numpy()
                for e in evs:
                    a, c = off[e], off[e + 1]
                    if mode == 0 and param > c - a:
                        assert n[e] == -1
                        continue
                    w = ref[e]
                    want = orc.exclusive_jets_def(p4[a:c], w[0], w[1], w[2], w[4], sch, mode, param)
                    assert n[e] == len(want) and o[a:a + n[e]].tobytes() == want.tobytes(), \
                        f"event {e} scheme {sch} mode {mode} param {param}"


@pytest.mark.gpu
def test_host_memory_batch(engine):
    """fjmi_cluster_batch_def from host arrays (pinned chunks): per-event
    definitions and radii with the history momenta, then one definition for all."""
    import fastjet_gpu_amd as fj
    p4, off = fj.gen_workload(3, 96, first=2000)
    ne = len(off) - 1
    ps = np.array([(0.5, -1.0, 1.0, -0.3, 0.0, 1.7)[e % 6] for e in range(ne)])
    schemes = np.array([e % 9 for e in range(ne)], np.int32)
    Rs = np.array([(0.4, 0.7, 1.0)[e % 3] for e in range(ne)])
    print(f"[gpu jetdef] host-memory batch x{ne} per-event definitions", flush=True)
    h1, h2, hd, jets, nj, hj = engine.cluster_batch_def(p4, off, ps, schemes, Rs, hist_jets=True)
    for e in range(ne):
        a, c = off[e], off[e + 1]
        w = orc.cluster_def_hist(p4[a:c], ps[e], int(schemes[e]), Rs[e])
        assert same(w[:4], (h1[a:c], h2[a:c], hd[a:c], jets[a:a + nj[e]])), f"event {e}"
        merge = w[1] >= 0
        assert w[4][merge].tobytes() == hj[a:c][merge].tobytes(), f"event {e} history momenta"
    print(f"[gpu jetdef] host-memory batch x{ne} one definition p=-1 WTA_modp R=0.5", flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch_def(p4, off, -1.0, fj.SCHEME_WTA_MODP, 0.5)
    for e in range(0, ne, 7):
        a, c = off[e], off[e + 1]
        w = orc.cluster_def(p4[a:c], -1.0, fj.SCHEME_WTA_MODP, 0.5)
        assert same(w, (h1[a:c], h2[a:c], hd[a:c], jets[a:a + nj[e]])), f"event {e}"
    bad = np.array([0.5] * ne)
    with pytest.raises(fj.FjmiError):
        engine.cluster_batch_def(p4, off, bad, np.array([0] * (ne - 1) + [12], np.int32), 0.4)
