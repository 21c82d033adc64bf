"""ctypes binding of the CPU oracle (oracle/_build/libfj_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product never imports it.
Build with `make -C oracle` (also done by `python -m fastjet_gpu_amd.build`).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_build", "libfj_oracle.so")
_lib = None
_vp = ctypes.c_void_p

MATH_SHARED, MATH_LIBM = 0, 1


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"{_PATH} missing; run `make -C oracle`")
        L = ctypes.CDLL(_PATH)
        for name in ("fjo_cluster_n2plain", "fjo_cluster_brute", "fjo_cluster_pairwise", "fjo_cluster_tiled"):
            f = getattr(L, name)
            f.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
            f.restype = ctypes.c_int
        for name in ("fjo_cluster_n2plain_def", "fjo_cluster_brute_def", "fjo_cluster_pairwise_def",
                     "fjo_cluster_tiled_def"):
            f = getattr(L, name)
            f.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp,
                          _vp, _vp, _vp]
            f.restype = ctypes.c_int
        for name in ("fjo_cluster_n2plain_defh", "fjo_cluster_brute_defh", "fjo_cluster_pairwise_defh",
                     "fjo_cluster_tiled_defh"):
            f = getattr(L, name)
            f.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _vp, _vp,
                          _vp, _vp, _vp, _vp]
            f.restype = ctypes.c_int
        L.fjo_exclusive_jets_def.argtypes = [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, _vp, _vp]
        L.fjo_exclusive_jets_def.restype = ctypes.c_int
        L.fjo_cluster_batch_def.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                            _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp]
        L.fjo_cluster_batch_def.restype = ctypes.c_double
        L.fjo_log.argtypes = [ctypes.c_double]
        L.fjo_log.restype = ctypes.c_double
        L.fjo_atan2.argtypes = [ctypes.c_double, ctypes.c_double]
        L.fjo_atan2.restype = ctypes.c_double
        for name, nargs in (("fjo_exp", 1), ("fjo_pow", 2), ("fjo_sqrt", 1), ("fjo_sin", 1), ("fjo_cos", 1),
                            ("fjo_weight_gen", 2)):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_double] * nargs
            f.restype = ctypes.c_double
        L.fjo_set_jet.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
        L.fjo_set_jet.restype = None
        L.fjo_time_batch.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.fjo_time_batch.restype = ctypes.c_double
        L.fjo_cluster_batch.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
        L.fjo_cluster_batch.restype = ctypes.c_double
        L.fjo_tiled_pair_evals.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.fjo_tiled_pair_evals.restype = ctypes.c_int64
        L.fjo_inclusive_jets.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _vp, _vp]
        L.fjo_inclusive_jets.restype = ctypes.c_int
        L.fjo_history_children.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp]
        L.fjo_history_children.restype = ctypes.c_int
        L.fjo_exclusive_jets.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _vp, _vp]
        L.fjo_exclusive_jets.restype = ctypes.c_int
        _lib = L
    return _lib


def _run(fname, p4, p, R, math_mode):
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    n = len(p4)
    h1 = np.zeros(max(n, 1), np.int32)
    h2 = np.zeros(max(n, 1), np.int32)
    hd = np.zeros(max(n, 1), np.float64)
    jets = np.zeros((max(n, 1), 4), np.float64)
    nj = np.zeros(1, np.int32)
    P = lambda a: a.ctypes.data_as(_vp)
    rc = getattr(lib(), fname)(P(p4), n, int(p), float(R), math_mode, P(h1), P(h2), P(hd), P(jets), P(nj))
    if rc != 0:
        raise ValueError(f"{fname} returned {rc}")
    return h1[:n], h2[:n], hd[:n], jets[:nj[0]].copy()


def _run_def(fname, p4, p, scheme, R):
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    n = len(p4)
    h1 = np.zeros(max(n, 1), np.int32)
    h2 = np.zeros(max(n, 1), np.int32)
    hd = np.zeros(max(n, 1), np.float64)
    jets = np.zeros((max(n, 1), 4), np.float64)
    nj = np.zeros(1, np.int32)
    rc = getattr(lib(), fname)(p4.ctypes.data_as(_vp), n, float(p), int(scheme), float(R), MATH_SHARED,
                               h1.ctypes.data_as(_vp), h2.ctypes.data_as(_vp), hd.ctypes.data_as(_vp),
                               jets.ctypes.data_as(_vp), nj.ctypes.data_as(_vp))
    if rc != 0:
        raise ValueError(f"{fname} failed ({rc})")
    return h1[:n], h2[:n], hd[:n], jets[:nj[0]]


def cluster_def(p4, p, scheme, R, method="n2plain"):
    """One event under a general jet definition: genkt exponent p (any finite real) and recombination scheme
    (0 E, 1 pt, 2 pt2, 3 Et, 4 Et2, 5 BIpt, 6 BIpt2, 7 WTA_pt, 8 WTA_modp).  Returns (h1, h2, dij, jets)."""
    return _run_def(f"fjo_cluster_{method}_def", p4, p, scheme, R)


def cluster_def_hist(p4, p, scheme, R, method="n2plain"):
    """cluster_def and the momentum of the history entry each merge step creates (N x 4; rows of
    beam steps are NaN).  Returns (h1, h2, dij, jets, hist_jets)."""
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    n = len(p4)
    h1 = np.zeros(max(n, 1), np.int32)
    h2 = np.zeros(max(n, 1), np.int32)
    hd = np.zeros(max(n, 1), np.float64)
    jets = np.zeros((max(n, 1), 4), np.float64)
    hj = np.full((max(n, 1), 4), np.nan)
    nj = np.zeros(1, np.int32)
    rc = getattr(lib(), f"fjo_cluster_{method}_defh")(_P(p4), n, float(p), int(scheme), float(R), MATH_SHARED, _P(h1),
                                                       _P(h2), _P(hd), _P(jets), _P(nj), _P(hj))
    if rc != 0:
        raise ValueError(f"fjo_cluster_{method}_defh failed ({rc})")
    return h1[:n], h2[:n], hd[:n], jets[:nj[0]].copy(), hj[:n]


def exclusive_jets_def(p4, h1, h2, hd, hist_jets, scheme, mode, param):
    """Exclusive jets of a clustering under recombination scheme `scheme`, from its stored history
    momenta (cluster_def_hist)."""
    p4 = np.ascontiguousarray(p4, np.float64).reshape(-1, 4)
    h1, h2 = np.ascontiguousarray(h1, np.int32), np.ascontiguousarray(h2, np.int32)
    hd = np.ascontiguousarray(hd, np.float64)
    hj = np.ascontiguousarray(hist_jets, np.float64).reshape(-1, 4)
    out = np.zeros((max(len(p4), 1), 4))
    k = np.zeros(1, np.int32)
    if lib().fjo_exclusive_jets_def(_P(p4), _P(h1), _P(h2), _P(hd), _P(hj), len(p4), int(scheme), int(mode),
                                    float(param), _P(out), _P(k)) != 0:
        raise ValueError("fjo_exclusive_jets_def failed")
    return out[:k[0]].copy()


def cluster_batch_def(p4, offsets, p, scheme, R, nthreads, method="tiled"):
    """A batch under one general jet definition (R scalar or per event).  Returns (h1, h2, hd, jets, nj)."""
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    ne, tot = len(offsets) - 1, len(p4)
    Ra = np.ascontiguousarray(R, dtype=np.float64) if np.ndim(R) else None
    h1 = np.zeros(max(tot, 1), np.int32)
    h2 = np.zeros(max(tot, 1), np.int32)
    hd = np.zeros(max(tot, 1), np.float64)
    jets = np.zeros((max(tot, 1), 4), np.float64)
    nj = np.zeros(max(ne, 1), np.int32)
    P = lambda a: None if a is None else a.ctypes.data_as(_vp)
    secs = lib().fjo_cluster_batch_def(P(p4), P(offsets), ne, float(p), int(scheme), 1.0 if Ra is not None else float(R),
                                       P(Ra), int(nthreads), 1 if method == "tiled" else 0, P(h1), P(h2), P(hd),
                                       P(jets), P(nj))
    if secs < 0:
        raise ValueError("fjo_cluster_batch_def failed")
    return h1[:tot], h2[:tot], hd[:tot], jets[:tot], nj[:ne]


def n2plain(p4, p, R, math_mode=MATH_SHARED):
    """FastJet N2Plain restatement: (hist_p1, hist_p2, hist_dij, jets)."""
    return _run("fjo_cluster_n2plain", p4, p, R, math_mode)


def brute(p4, p, R, math_mode=MATH_SHARED):
    return _run("fjo_cluster_brute", p4, p, R, math_mode)


def pairwise(p4, p, R, math_mode=MATH_SHARED):
    return _run("fjo_cluster_pairwise", p4, p, R, math_mode)


def tiled(p4, p, R, math_mode=MATH_SHARED):
    """N2Plain decisions with tiled nearest-neighbour bookkeeping."""
    return _run("fjo_cluster_tiled", p4, p, R, math_mode)


def time_batch(p4, offsets, p, R, nthreads, method="tiled"):
    """Wall seconds to cluster the events on `nthreads` host threads."""
    p4 = np.ascontiguousarray(p4, dtype=np.float64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    return lib().fjo_time_batch(p4.ctypes.data_as(_vp), offsets.ctypes.data_as(_vp), len(offsets) - 1, int(p),
                                float(R), int(nthreads), 1 if method == "tiled" else 0)


def cluster_batch(p4, offsets, p, R, nthreads, method="tiled", math_mode=MATH_SHARED):
    """Clusters every event of a packed batch on `nthreads` host threads; p and R
    are scalars or per-event arrays.  Returns (hist_p1, hist_p2, hist_dij, jets,
    njets, pair_evals, seconds) with per-particle arrays indexed like the input
    and pair_evals per event (tiled method; -1 for n2plain)."""
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    ne, tot = len(offsets) - 1, len(p4)
    pa = np.ascontiguousarray(p, dtype=np.int32) if np.ndim(p) else None
    Ra = np.ascontiguousarray(R, dtype=np.float64) if np.ndim(R) else None
    h1 = np.zeros(max(tot, 1), np.int32)
    h2 = np.zeros(max(tot, 1), np.int32)
    hd = np.zeros(max(tot, 1), np.float64)
    jets = np.zeros((max(tot, 1), 4), np.float64)
    nj = np.zeros(max(ne, 1), np.int32)
    pairs = np.zeros(max(ne, 1), np.int64)
    P = lambda a: None if a is None else a.ctypes.data_as(_vp)
    secs = lib().fjo_cluster_batch(P(p4), P(offsets), ne, 0 if pa is not None else int(p),
                                   1.0 if Ra is not None else float(R), P(pa), P(Ra), int(nthreads),
                                   1 if method == "tiled" else 0, math_mode, P(h1), P(h2), P(hd), P(jets), P(nj),
                                   P(pairs))
    if secs < 0:
        raise ValueError("fjo_cluster_batch failed")
    return h1[:tot], h2[:tot], hd[:tot], jets[:tot], nj[:ne], pairs[:ne], secs


def tiled_pair_evals(p4, p, R):
    """Pair distances the tiled restatement evaluates for one event."""
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    r = lib().fjo_tiled_pair_evals(p4.ctypes.data_as(_vp), len(p4), int(p), float(R))
    if r < 0:
        raise ValueError("fjo_tiled_pair_evals failed")
    return int(r)


def _P(a):
    return a.ctypes.data_as(_vp)


def inclusive_jets(jets, ptmin):
    """Jets with pt^2 >= ptmin^2 sorted by decreasing pt^2 (main.cc:309-327)."""
    jets = np.ascontiguousarray(jets, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((max(len(jets), 1), 4))
    n = np.zeros(1, np.int32)
    if lib().fjo_inclusive_jets(_P(jets), len(jets), float(ptmin), _P(out), _P(n)) != 0:
        raise ValueError("fjo_inclusive_jets failed")
    return out[:n[0]].copy()


def history_children(h1, h2, hd):
    """(child[2N], max_dij[N]) of a history (FastJet's ClusterSequence view)."""
    h1, h2 = np.ascontiguousarray(h1, np.int32), np.ascontiguousarray(h2, np.int32)
    hd = np.ascontiguousarray(hd, np.float64)
    n = len(h1)
    child = np.zeros(max(2 * n, 1), np.int32)
    md = np.zeros(max(n, 1), np.float64)
    if lib().fjo_history_children(_P(h1), _P(h2), _P(hd), n, _P(child), _P(md)) != 0:
        raise ValueError("fjo_history_children failed")
    return child[:2 * n], md[:n]


def exclusive_jets(p4, h1, h2, hd, mode, param):
    """FastJet's exclusive_jets(njets) (mode 0) or exclusive_jets(dcut) (mode 1) from a history."""
    p4 = np.ascontiguousarray(p4, np.float64).reshape(-1, 4)
    h1, h2 = np.ascontiguousarray(h1, np.int32), np.ascontiguousarray(h2, np.int32)
    hd = np.ascontiguousarray(hd, np.float64)
    out = np.zeros((max(len(p4), 1), 4))
    k = np.zeros(1, np.int32)
    if lib().fjo_exclusive_jets(_P(p4), _P(h1), _P(h2), _P(hd), len(p4), int(mode), float(param), _P(out), _P(k)) != 0:
        raise ValueError("fjo_exclusive_jets failed")
    return out[:k[0]].copy()


def set_jet(p4row, p, math_mode=MATH_SHARED):
    q = np.ascontiguousarray(p4row, dtype=np.float64)
    r, f, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().fjo_set_jet(q.ctypes.data_as(_vp), int(p), math_mode, ctypes.byref(r), ctypes.byref(f), ctypes.byref(w))
    return r.value, f.value, w.value


def read_events(path):
    """Events of a reference-format text file (main.cc:34-70): 'px py pz E' lines;
    a '#' or blank line ends an event."""
    events, cur = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 4 and not line.startswith("#"):
                try:
                    cur.append([float(x) for x in parts[:4]])
                    continue
                except ValueError:
                    pass
            if cur:
                events.append(np.array(cur, np.float64))
                cur = []
    if cur:
        events.append(np.array(cur, np.float64))
    return events
