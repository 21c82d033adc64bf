"""Pins of the oracle's history-derived results (oracle/fj_oracle.c), which the
GPU's post-processing kernels are checked against:

  inclusive jets  permutation of the jets above ptmin, in non-increasing pt^2,
                  none below ptmin (main.cc:309-327);
  history view    each entry consumed at most once, by a later step that names
                  it as a parent; the unconsumed entries are exactly the beam
                  steps' results; max_dij is the running maximum;
  exclusive jets  njets = N gives the particles themselves; from n + 1 to n jets
                  the step at the stop point either merges two of them (their
                  sum, bit for bit) or sends one to the beam; the count is n;
                  dcut selects the stop point after the last step whose running
                  maximum is <= dcut, and counts do not grow with dcut.
Checked on the reference's input files and on generated events; no formula of
the oracle is restated here.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.conftest import GOLDEN

ALGOS = [1, 0, -1]


def _events():
    evs = orc.read_events(os.path.join(GOLDEN, "multi-event.dat"))[:4]
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 40):
        pt = 0.5 + rng.exponential(2.0, n)
        eta = rng.uniform(-2, 2, n)
        phi = rng.uniform(0, 2 * np.pi, n)
        evs.append(np.stack([pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta), pt * np.cosh(eta)], 1))
    return evs


def _rows(a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)
    return sorted(a[k].tobytes() for k in range(len(a)))


@pytest.mark.parametrize("p", ALGOS)
def test_inclusive_jets(p):
    for ev in _events():
        jets = orc.n2plain(ev, p, 0.6)[3]
        for ptmin in (0.0, 0.7, 2.5, 1e9):
            got = orc.inclusive_jets(jets, ptmin)
            key = jets[:, 0] * jets[:, 0] + jets[:, 1] * jets[:, 1]
            assert _rows(got) == _rows(jets[~(key < ptmin * ptmin)])
            gk = got[:, 0] * got[:, 0] + got[:, 1] * got[:, 1]
            assert np.all(gk[:-1] >= gk[1:]) and np.all(gk >= ptmin * ptmin)


@pytest.mark.parametrize("p", ALGOS)
def test_history_view(p):
    for ev in _events():
        n = len(ev)
        h1, h2, hd, jets = orc.n2plain(ev, p, 0.6)
        child, md = orc.history_children(h1, h2, hd)
        consumed = {}
        for s in range(n):
            for par in (h1[s], h2[s]):
                if par >= 0:
                    assert par not in consumed
                    consumed[par] = n + s
        for k in range(2 * n):
            assert child[k] == consumed.get(k, -3)
            if child[k] >= 0:
                assert child[k] > k
        beam_results = {n + s for s in range(n) if h2[s] < 0}
        assert {k for k in range(2 * n) if child[k] == -3} == beam_results
        assert len(beam_results) == len(jets)
        assert np.all(md >= hd) and np.all(np.diff(md) >= 0)
        assert np.all([md[s] in (hd[:s + 1]) for s in range(n)])


@pytest.mark.parametrize("p", ALGOS)
def test_exclusive_jets_by_count(p):
    for ev in _events():
        n = len(ev)
        h1, h2, hd, _ = orc.n2plain(ev, p, 0.6)
        all_ = orc.exclusive_jets(ev, h1, h2, hd, 0, n)
        assert _rows(all_) == _rows(ev)  # every particle, unchanged
        assert len(orc.exclusive_jets(ev, h1, h2, hd, 0, 0)) == 0
        prev = all_
        for nj in range(n - 1, -1, -1):
            cur = orc.exclusive_jets(ev, h1, h2, hd, 0, nj)
            assert len(cur) == nj
            s = n - nj - 1  # the step at the stop point of nj + 1 jets
            pr, cr = _rows(prev), _rows(cur)
            if h2[s] >= 0:
                gone = [r for r in pr if r not in cr]
                new = [r for r in cr if r not in pr]
                assert len(gone) == 2 and len(new) == 1
                a, b = (np.frombuffer(g, np.float64) for g in gone)
                assert np.frombuffer(new[0], np.float64).tobytes() == (a + b).tobytes()
            else:
                assert len(pr) == len(cr) + 1 and all(r in pr for r in cr)
            prev = cur
        with pytest.raises(ValueError):
            orc.exclusive_jets(ev, h1, h2, hd, 0, n + 1)


@pytest.mark.parametrize("p", ALGOS)
def test_exclusive_jets_by_dcut(p):
    for ev in _events():
        n = len(ev)
        h1, h2, hd, _ = orc.n2plain(ev, p, 0.6)
        _, md = orc.history_children(h1, h2, hd)
        last = None
        for dcut in sorted({0.0, *md.tolist(), *(0.5 * md).tolist(), float(md[-1]) * 2 + 1.0}):
            got = orc.exclusive_jets(ev, h1, h2, hd, 1, dcut)
            first_over = next((s for s in range(n) if md[s] > dcut), n)
            want = orc.exclusive_jets(ev, h1, h2, hd, 0, n - first_over)
            assert got.tobytes() == want.tobytes()
            if last is not None:
                assert len(got) <= last
            last = len(got)
        assert len(orc.exclusive_jets(ev, h1, h2, hd, 1, float(md[-1]))) == 0
