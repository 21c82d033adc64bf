"""Pins the CPU oracle (oracle/fj_oracle.c, FastJet N2Plain restatement).

The reference ships no expected outputs (its output/ logs are git-ignored), so the
oracle is pinned by checks that a plausible mistake would fail:
- brute force (every nearest neighbour recomputed from scratch each step) and the
  textbook all-pairs generalised-kt definition give the same history;
- invariants: every input ends in exactly one jet, four-momentum conservation,
  exclusivity of the final C/A jets (pairwise dR >= R: all merges precede the beam steps);
- closed-form cases: isolated particles, two-particle events, C/A beam order;
- the committed golden hashes of its output on the reference's input files;
- the glibc-math mode (what FastJet itself uses) gives the same merge sequence.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

from oracle import oracle as orc
from tests.conftest import GOLDEN

ALGOS = [1, 0, -1]


def random_event(rng, n, spread=1.0):
    pt = 0.5 + rng.exponential(5.0, n)
    eta = rng.uniform(-2.0 * spread, 2.0 * spread, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    m = rng.uniform(0, 0.3, n)
    px, py, pz = pt * np.cos(phi), pt * np.sin(phi), pt * np.sinh(eta)
    E = np.sqrt(px * px + py * py + pz * pz + m * m)
    return np.stack([px, py, pz, E], axis=1)


def same(a, b):
    return all(np.asarray(x).tobytes() == np.asarray(y).tobytes() for x, y in zip(a, b))


@settings(max_examples=60, deadline=None)
@given(n=st.integers(1, 40), seed=st.integers(0, 2 ** 31 - 1), p=st.sampled_from(ALGOS),
       R=st.sampled_from([0.2, 0.4, 0.7, 1.0, 1.5]))
def test_n2plain_equals_brute_force(n, seed, p, R):
    ev = random_event(np.random.default_rng(seed), n)
    assert same(orc.n2plain(ev, p, R), orc.brute(ev, p, R))


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("p", [1, -1])
def test_n2plain_equals_textbook_pairwise(seed, p):
    ev = random_event(np.random.default_rng(100 + seed), 150)
    assert same(orc.n2plain(ev, p, 0.6), orc.pairwise(ev, p, 0.6))


@pytest.mark.parametrize("seed", range(4))
def test_cambridge_matches_pairwise_up_to_tied_beams(seed):
    """C/A: every beam distance is R^2, so the merges agree exactly and the jets
    agree as a set (the order of the tied beam steps is a slot convention)."""
    ev = random_event(np.random.default_rng(200 + seed), 150)
    a, b = orc.n2plain(ev, 0, 0.6), orc.pairwise(ev, 0, 0.6)
    merges = lambda r: [(x, y) for x, y in zip(r[0], r[1]) if y >= 0]
    assert merges(a) == merges(b)
    key = lambda j: sorted(map(tuple, np.asarray(j).tolist()))
    assert key(a[3]) == key(b[3])


def walk(ev, res):
    """Checks the history structure and returns the constituents of each jet."""
    h1, h2, hd, jets = res
    n = len(ev)
    assert len(h1) == n
    members = {i: [i] for i in range(n)}
    used = set()
    out = []
    for s in range(n):
        a, b = int(h1[s]), int(h2[s])
        assert a not in used and 0 <= a < n + s
        used.add(a)
        if b >= 0:
            assert a < b and b not in used and b < n + s
            used.add(b)
            members[n + s] = members.pop(a) + members.pop(b)
        else:
            out.append(members.pop(a))
    assert not members and len(out) == len(jets)
    return out


@pytest.mark.parametrize("p", ALGOS)
@pytest.mark.parametrize("R", [0.4, 1.0])
def test_invariants_on_reference_file(p, R):
    ev = orc.read_events(os.path.join(GOLDEN, "808_cms.dat"))[0]
    res = orc.n2plain(ev, p, R)
    groups = walk(ev, res)
    assert sorted(i for g in groups for i in g) == list(range(len(ev)))      # each input in one jet
    for g, j in zip(groups, res[3]):
        assert np.allclose(ev[g].sum(axis=0), j, rtol=1e-12, atol=1e-9)      # E-scheme = 4-vector sum
    tot = np.abs(res[3]).sum(axis=0)
    assert np.allclose(res[3].sum(axis=0), ev.sum(axis=0), rtol=0, atol=1e-12 * tot.max())
    if p == 0:   # C/A: every merge precedes every beam step, so final jets are >= R apart
        raps = [orc.set_jet(j, p)[:2] for j in res[3]]
        for a in range(len(raps)):
            for b in range(a + 1, len(raps)):
                dphi = abs(raps[a][1] - raps[b][1])
                dphi = min(dphi, 2 * math.pi - dphi)
                assert dphi ** 2 + (raps[a][0] - raps[b][0]) ** 2 >= R * R * (1 - 1e-12)


def test_isolated_particles_beam_order():
    """Particles pairwise farther apart than R never merge; kt promotes the softest
    first, anti-kt the hardest first, C/A (all tied) by slot: 0, N-1, N-2, ..., 1."""
    n = 12
    pts = np.array([3.0, 1.0, 7.0, 2.0, 9.0, 4.0, 8.0, 5.0, 6.0, 11.0, 10.0, 12.0])
    phi = np.arange(n) * (2 * math.pi / n)
    ev = np.stack([pts * np.cos(phi), pts * np.sin(phi), np.zeros(n), pts], axis=1)
    for p, order in ((1, np.argsort(pts)), (-1, np.argsort(-pts)), (0, [0] + list(range(n - 1, 0, -1)))):
        h1, h2, hd, jets = orc.n2plain(ev, p, 0.4)
        assert np.all(h2 == -1)
        assert list(h1) == list(order)
        assert np.array_equal(jets, ev[list(order)])


@pytest.mark.parametrize("dr", [0.1, 0.39, 0.41, 0.9])
@pytest.mark.parametrize("p", ALGOS)
def test_two_particles(dr, p):
    a = np.array([10.0, 0.0, 0.0, 10.0])
    b = np.array([3.0 * math.cos(dr), 3.0 * math.sin(dr), 0.0, 3.0])
    h1, h2, hd, jets = orc.n2plain(np.stack([a, b]), p, 0.4)
    if dr < 0.4:
        assert (h1[0], h2[0]) == (0, 1) and len(jets) == 1
        assert np.array_equal(jets[0], a + b)
        w = {1: 9.0, 0: 1.0, -1: 1.0 / 100.0}[p]
        assert hd[0] == pytest.approx(w * dr * dr / 0.16, rel=1e-12)
    else:
        assert np.all(h2 == -1) and len(jets) == 2


def result_hash(res):
    h = hashlib.sha256()
    for x in res:
        h.update(np.ascontiguousarray(x).tobytes())
    return h.hexdigest()


def test_golden_hashes():
    """Committed hashes of the oracle's output on the reference's input files
    (written by tests/golden/make_golden.py)."""
    with open(os.path.join(GOLDEN, "oracle_hashes.json")) as fh:
        gold = json.load(fh)
    assert gold
    for key, want in gold.items():
        fname, idx, p, R = key.split("|")
        ev = orc.read_events(os.path.join(GOLDEN, fname))[int(idx)]
        assert result_hash(orc.n2plain(ev, int(p), float(R))) == want, key


@pytest.mark.parametrize("fname", ["354_fastjet.dat", "808_cms.dat", "multi-event.dat"])
def test_glibc_math_gives_same_merges(fname):
    """FastJet evaluates rap/phi with glibc; the shared correctly rounded functions
    must not change any clustering decision on the reference's inputs."""
    for ev in orc.read_events(os.path.join(GOLDEN, fname)):
        for p in ALGOS:
            for R in (0.4, 1.0):
                a = orc.n2plain(ev, p, R, orc.MATH_SHARED)
                b = orc.n2plain(ev, p, R, orc.MATH_LIBM)
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
                assert np.allclose(a[2], b[2], rtol=1e-12, atol=0)
                assert np.array_equal(a[3], b[3])


@settings(max_examples=60, deadline=None)
@given(n=st.integers(1, 120), seed=st.integers(0, 2 ** 31 - 1), p=st.sampled_from(ALGOS),
       R=st.sampled_from([0.1, 0.4, 0.7, 1.0, 2.5, 7.0]), spread=st.sampled_from([0.2, 1.0, 3.0]))
def test_tiled_equals_n2plain(n, seed, p, R, spread):
    """The tiled bookkeeping gives the N2Plain decisions bit for bit, including
    R > 2pi/3 (fewer than three phi tiles) and very dense or sparse events."""
    ev = random_event(np.random.default_rng(seed), n, spread)
    assert same(orc.n2plain(ev, p, R), orc.tiled(ev, p, R))


@pytest.mark.parametrize("p", ALGOS)
def test_tiled_equals_n2plain_on_ties(p):
    """Exact duplicates and a regular lattice: every kind of exact tie."""
    base = np.array([[3.0, 1.0, 0.5, 4.0], [0.5, 2.0, -1.0, 3.0], [-1.0, -1.0, 0.2, 2.0]])
    dup = np.repeat(base, 30, axis=0)
    g = []
    for i in range(14):
        for j in range(14):
            eta, phi = -1.5 + 0.25 * i, 0.25 * j
            pt = 1.0 + 0.125 * ((i * 7 + j * 3) % 5)
            g.append([pt * math.cos(phi), pt * math.sin(phi), pt * math.sinh(eta), pt * math.cosh(eta)])
    for ev, R in ((dup, 0.7), (np.array(g), 0.3), (np.array(g), 0.25), (np.array(g), 0.5)):
        assert same(orc.n2plain(ev, p, R), orc.tiled(ev, p, R))


def test_tiled_equals_n2plain_on_reference_files():
    for fname in ("354_fastjet.dat", "808_cms.dat", "multi-event.dat"):
        for ev in orc.read_events(os.path.join(GOLDEN, fname)):
            for p in ALGOS:
                for R in (0.4, 1.0, 3.0):
                    assert same(orc.n2plain(ev, p, R), orc.tiled(ev, p, R)), (fname, len(ev), p, R)
