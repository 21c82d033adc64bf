"""GPU parity at the tiled engine's limits: the largest events it takes (N up
to 65535, beyond the 20000 of the high-pileup workload), radii of 2 pi and
more (a single phi column that is its own neighbour on both sides; the
reference divides by zero there, grid.cu:72), and the refusal of larger events.
Bar: bit-exact history and jets against the oracle's tiled restatement (pinned
to N2Plain in tests/test_oracle*.py)."""
import numpy as np
import pytest

from oracle import oracle as orc
from tests.test_gpu_parity import first_diff, same


def _pileup(n):
    """The first n particles of high-pileup events 0, 1, 2, ... concatenated."""
    import fastjet_gpu_amd as fj
    evs, e = [], 0
    while sum(len(x) for x in evs) < n:
        evs.append(fj.gen_event(4, e))
        e += 1
    return np.ascontiguousarray(np.concatenate(evs)[:n])


@pytest.mark.gpu
@pytest.mark.parametrize("n,p", [(40000, 1), (40000, -1), (65535, -1)])
def test_largest_tiled_events(engine, n, p):
    ev = _pileup(n)
    print(f"[gpu limits] N={n} p={p} R=0.4 tiled", flush=True)
    got = engine.cluster_event(ev, p, 0.4, engine=2)
    want = orc.tiled(ev, p, 0.4)
    assert same(want, got), f"N={n}: {first_diff(want, got)}"


@pytest.mark.gpu
def test_largest_tiled_event_in_a_batch(engine):
    """N = 65535 beside small events in one graph launch (AUTO buckets)."""
    import fastjet_gpu_amd as fj
    big = _pileup(65535)
    small, soff = fj.gen_workload(3, 6, first=900)
    p4 = np.concatenate([small[:soff[3]], big, small[soff[3]:]])
    sizes = [*np.diff(soff)[:3], len(big), *np.diff(soff)[3:]]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    print(f"[gpu limits] batch with N={sizes}", flush=True)
    h1, h2, hd, jets, nj = engine.cluster_batch(p4, off, -1, 0.4)
    for e in range(len(sizes)):
        a, b = off[e], off[e + 1]
        want = orc.tiled(p4[a:b], -1, 0.4)
        got = (h1[a:b], h2[a:b], hd[a:b], jets[a:a + nj[e]])
        assert same(want, got), f"event {e} (N={b - a}): {first_diff(want, got)}"


@pytest.mark.gpu
@pytest.mark.parametrize("R", [2 * np.pi, 6.3, 7.0, 12.0])
@pytest.mark.parametrize("p", [1, 0, -1])
def test_radius_beyond_two_pi(engine, R, p):
    """One phi column (side 2 pi), wide rapidity spread so that several jets
    remain; dense and tiled engines."""
    rng = np.random.default_rng(31)
    n = 900
    pt = 0.3 + rng.exponential(1.0, n)
    eta = rng.uniform(-30.0, 30.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    ev = np.stack([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)], 1)
    want = orc.n2plain(ev, p, R)
    assert len(want[3]) > 1
    for eng in (1, 2):
        print(f"[gpu limits] R={R} p={p} N={n} engine={eng}", flush=True)
        got = engine.cluster_event(ev, p, R, engine=eng)
        assert same(want, got), f"R={R} engine {eng}: {first_diff(want, got)}"


@pytest.mark.gpu
def test_events_beyond_the_tiled_limit_are_refused(engine):
    """N > 65535: FJMI_ECAPACITY from the one-event and batch entry points,
    before anything is launched; the context stays usable."""
    import fastjet_gpu_amd as fj
    ev = _pileup(65536)
    for eng in (0, 2):
        with pytest.raises(fj.FjmiError, match=r"\(-3\)"):
            engine.cluster_event(ev, -1, 0.4, engine=eng)
    off = np.array([0, 100, 100 + len(ev)], np.int64)
    p4 = np.concatenate([ev[:100], ev])
    with pytest.raises(fj.FjmiError, match=r"\(-3\)"):
        engine.cluster_batch(p4, off, -1, 0.4)
    small = ev[:500]
    assert same(orc.n2plain(small, -1, 0.4), engine.cluster_event(small, -1, 0.4))
