"""fjmi_cluster_event_device_noalloc (through the C ABI): the one-event entry
point that allocates nothing, its scratch a static array of the library.  Its
results must equal fjmi_cluster_event_device's bit for bit, on two calls in a
row where the second event is smaller (the array keeps the first call's
contents), and events beyond its stated maximum are refused."""
import numpy as np
import pytest

from oracle import oracle as orc


def _bufs(torch, n):
    return dict(h1=torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                h2=torch.full((n,), -7, dtype=torch.int32, device="cuda"),
                hd=torch.zeros(n, dtype=torch.float64, device="cuda"),
                jets=torch.zeros(4 * n, dtype=torch.float64, device="cuda"),
                nj=torch.full((1,), -2, dtype=torch.int32, device="cuda"))


def _host(b):
    nj = int(b["nj"].cpu()[0])
    return (b["h1"].cpu().numpy(), b["h2"].cpu().numpy(), b["hd"].cpu().numpy(),
            b["jets"].cpu().numpy().reshape(-1, 4)[:max(nj, 0)], nj)


def _args(b):
    return (b["h1"].data_ptr(), b["h2"].data_ptr(), b["hd"].data_ptr(), b["jets"].data_ptr(), b["nj"].data_ptr())


@pytest.mark.gpu
def test_noalloc_matches_device_entry_point(engine):
    import torch
    import fastjet_gpu_amd as fj
    big = fj.gen_event(4, 1)[:12000]        # first call: large
    small = fj.gen_event(3, 3)              # second call: smaller, same static array
    tiny = fj.gen_event(1, 0)[:40]
    s = torch.cuda.Stream()
    for k, (ev, p, R) in enumerate([(big, -1, 0.4), (small, 1, 0.4), (tiny, 0, 1.0)]):
        n = len(ev)
        d = torch.from_numpy(np.ascontiguousarray(ev)).cuda()
        ref, got = _bufs(torch, n), _bufs(torch, n)
        torch.cuda.synchronize()
        print(f"[no-alloc] call {k}: N={n} p={p} R={R}", flush=True)
        engine.cluster_event_device(d.data_ptr(), n, p, R, *_args(ref), s.cuda_stream, engine=fj.ENGINE_TILED)
        engine.stream_check(s.cuda_stream)
        engine.cluster_event_device_noalloc(d.data_ptr(), n, p, R, *_args(got), s.cuda_stream)
        s.synchronize()
        r, g = _host(ref), _host(got)
        assert g[4] == r[4] and g[4] > 0
        for x, y in zip(r[:4], g[:4]):
            assert np.asarray(x).tobytes() == np.asarray(y).tobytes(), f"call {k}: N={n}"
        if n <= 3000:
            want = orc.n2plain(ev, p, R)
            for x, y in zip(want, g[:4]):
                assert np.asarray(x).tobytes() == np.asarray(y).tobytes()


@pytest.mark.gpu
def test_noalloc_refuses_events_beyond_its_maximum(engine):
    import torch
    import fastjet_gpu_amd as fj
    ev = np.concatenate([fj.gen_event(4, 2), fj.gen_event(4, 0)])[:20001]
    assert len(ev) == 20001
    d = torch.from_numpy(np.ascontiguousarray(ev)).cuda()
    b = _bufs(torch, len(ev))
    torch.cuda.synchronize()
    with pytest.raises(fj.FjmiError, match=r"\(-3\).*at most 20000"):
        engine.cluster_event_device_noalloc(d.data_ptr(), len(ev), -1, 0.4, *_args(b), 0)
    assert int(b["nj"].cpu()[0]) == -2  # nothing was launched


@pytest.mark.gpu
def test_noalloc_empty_event(engine):
    import torch
    b = _bufs(torch, 1)
    torch.cuda.synchronize()
    engine.cluster_event_device_noalloc(0, 0, -1, 0.4, *_args(b), 0)
    torch.cuda.synchronize()
    assert int(b["nj"].cpu()[0]) == 0
