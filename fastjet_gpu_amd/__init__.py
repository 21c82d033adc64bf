"""fastjet_gpu_amd — MI355X (gfx950) sequential-recombination jet clustering.

Python binding (ctypes) over the C ABI of lib/libfjmi.so (include/fjmi.h) and the
event generator lib/libfjmi_gen.so (include/fjmi_gen.h).  The clustering itself
runs only in the HIP kernels: if the library is missing, importing the engine
raises instead of falling back to anything else.

Algorithm exponents follow the reference's Algorithm enum (cluster.h:4):
kt = 1, Cambridge/Aachen = 0, anti-kt = -1.
"""
import ctypes
import os

import numpy as np

__all__ = ["Engine", "Server", "FjmiError", "KT", "CAMBRIDGE_AACHEN", "ANTIKT", "ENGINE_AUTO", "ENGINE_DENSE",
           "ENGINE_TILED", "ENGINE_EMPTY", "gen_event", "gen_serving_attrs", "gen_workload", "lib_path", "load_lib",
           "write_events", "read_events", "read_text_events", "JetDef", "JETDEF_DTYPE", "SCHEME_E", "SCHEME_PT",
           "SCHEME_PT2", "SCHEME_ET", "SCHEME_ET2", "SCHEME_BIPT", "SCHEME_BIPT2", "SCHEME_WTA_PT", "SCHEME_WTA_MODP",
           "NUM_SCHEMES"]

KT, CAMBRIDGE_AACHEN, ANTIKT = 1, 0, -1
# recombination schemes of general jet definitions (FastJet's RecombinationScheme numbers, include/fjmi.h)
SCHEME_E, SCHEME_PT, SCHEME_PT2, SCHEME_ET, SCHEME_ET2, SCHEME_BIPT, SCHEME_BIPT2, SCHEME_WTA_PT, SCHEME_WTA_MODP = \
    range(9)
NUM_SCHEMES = 9
EXCLUSIVE_NJETS, EXCLUSIVE_DCUT = 0, 1
ENGINE_AUTO, ENGINE_DENSE, ENGINE_TILED, ENGINE_EMPTY = 0, 1, 2, 3

_PKG = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_PKG, "lib")

_c_int_p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


class FjmiError(RuntimeError):
    pass


FJMI_EAGAIN = -6


class Bucket(ctypes.Structure):
    _fields_ = [("max_n", ctypes.c_int32), ("engine", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("slots", ctypes.c_int32), ("p4_global", ctypes.c_int32)]


class JetDef(ctypes.Structure):
    """fjmi_jetdef: genkt exponent p (any finite real) and recombination scheme SCHEME_*."""
    _fields_ = [("p", ctypes.c_double), ("scheme", ctypes.c_int32), ("pad", ctypes.c_int32)]


# fjmi_jetdef as a numpy record (per-event definitions in device memory)
JETDEF_DTYPE = np.dtype([("p", np.float64), ("scheme", np.int32), ("pad", np.int32)])


class BucketStats(ctypes.Structure):
    _fields_ = [("bucket", Bucket), ("n_events", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("particles", ctypes.c_int64), ("sum_n2", ctypes.c_double), ("ms", ctypes.c_float)]


def lib_path(name="libfjmi.so"):
    return os.path.join(_LIB, name)


_libs = {}


def load_lib(name="libfjmi.so"):
    """Loads a native library of the package; raises if it has not been built."""
    if name in _libs:
        return _libs[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise FjmiError(f"{path} is not built; run `python -m fastjet_gpu_amd.build`")
    lib = ctypes.CDLL(path)
    if name == "libfjmi.so":
        _declare(lib)
    elif name == "libfjmi_gen.so":
        lib.fjmi_gen_event_size.argtypes = [ctypes.c_int, ctypes.c_int64]
        lib.fjmi_gen_event_size.restype = ctypes.c_int
        lib.fjmi_gen_event.argtypes = [ctypes.c_int, ctypes.c_int64, _vp, ctypes.c_int]
        lib.fjmi_gen_event.restype = ctypes.c_int
        lib.fjmi_gen_serving_attrs.argtypes = [ctypes.c_int64, _c_int_p, ctypes.POINTER(ctypes.c_double),
                                               ctypes.POINTER(ctypes.c_double)]
        lib.fjmi_gen_serving_attrs.restype = ctypes.c_int
    _libs[name] = lib
    return lib


class ServerConfig(ctypes.Structure):
    _fields_ = [("device_budget", ctypes.c_int64), ("host_budget", ctypes.c_int64), ("max_events", ctypes.c_int32),
                ("variant", ctypes.c_int32), ("tile_q", ctypes.c_int32), ("idle_us", ctypes.c_int32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("handle", ctypes.c_int32), ("n", ctypes.c_int32),
                ("njets", ctypes.c_int32), ("pad", ctypes.c_int32), ("hist_p1", _c_int_p), ("hist_p2", _c_int_p),
                ("hist_dij", ctypes.POINTER(ctypes.c_double)), ("jets", ctypes.POINTER(ctypes.c_double))]


class ServerInfo(ctypes.Structure):
    _fields_ = [("particle_capacity", ctypes.c_int64), ("event_capacity", ctypes.c_int32), ("grid", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("host_bytes", ctypes.c_int64), ("submitted", ctypes.c_int64),
                ("completed", ctypes.c_int64), ("in_flight", ctypes.c_int64),
                ("particles_in_flight", ctypes.c_int64), ("peak_particles_in_flight", ctypes.c_int64),
                ("launches", ctypes.c_int64), ("gpu_peak_busy", ctypes.c_int64)]


def _declare(lib):
    d = ctypes.c_double
    i = ctypes.c_int
    sig = {
        "fjmi_last_error": ([], ctypes.c_char_p),
        "fjmi_version": ([], i),
        "fjmi_create": ([i, ctypes.POINTER(_vp)], i),
        "fjmi_destroy": ([_vp], i),
        "fjmi_kernel_configs": ([_vp, i], i),
        "fjmi_set_buckets": ([_vp, _vp, i], i),
        "fjmi_get_buckets": ([_vp, _vp, i, _c_int_p], i),
        "fjmi_cluster_event": ([_vp, _vp, i, i, d, i, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_cluster_event_device": ([_vp, _vp, i, i, d, i, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_cluster_pseudojets": ([_vp, _vp, i, i, d, i, _vp], i),
        "fjmi_stream_check": ([_vp, _vp], i),
        "fjmi_events_write": ([ctypes.c_char_p, i, _vp, _vp, _vp, _vp], i),
        "fjmi_events_info": ([ctypes.c_char_p, _c_int_p, ctypes.POINTER(ctypes.c_int64), _c_int_p, _c_int_p], i),
        "fjmi_events_read": ([ctypes.c_char_p, _vp, _vp, _vp, _vp], i),
        "fjmi_inclusive_jets": ([_vp, i, _vp, _vp, _vp, d, _vp, _vp, _vp], i),
        "fjmi_inclusive_pseudojets": ([_vp, _vp, i, d, _vp, _vp, _vp], i),
        "fjmi_history_view": ([_vp, i, _vp, i, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_exclusive_jets": ([_vp, i, _vp, i, _vp, _vp, _vp, _vp, i, d, _vp, _vp, _vp, _vp], i),
        "fjmi_cluster_event_device_noalloc": ([_vp, _vp, i, i, d, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_batch_create": ([_vp, i, _vp, _vp, _vp, _vp, _vp, i, d, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)],
                              i),
        "fjmi_cluster_event_def": ([_vp, _vp, i, _vp, d, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_cluster_event_device_def": ([_vp, _vp, i, _vp, d, _vp, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_batch_create_def": ([_vp, i, _vp, _vp, _vp, _vp, _vp, _vp, d, _vp, _vp, _vp, _vp, _vp, _vp,
                                   ctypes.POINTER(_vp)], i),
        "fjmi_cluster_batch_def": ([_vp, i, _vp, _vp, _vp, _vp, _vp, d, _vp, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_exclusive_jets_def": ([_vp, i, _vp, i, _vp, _vp, _vp, _vp, _vp, i, i, d, _vp, _vp, _vp], i),
        "fjmi_batch_launch": ([_vp, _vp], i),
        "fjmi_batch_sync": ([_vp, _vp], i),
        "fjmi_batch_destroy": ([_vp], i),
        "fjmi_batch_stats": ([_vp, _vp, i, _c_int_p], i),
        "fjmi_cluster_batch": ([_vp, i, _vp, _vp, _vp, _vp, i, d, _vp, _vp, _vp, _vp, _vp], i),
        "fjmi_server_create": ([_vp, _vp, ctypes.POINTER(_vp)], i),
        "fjmi_server_submit": ([_vp, _vp, i, i, d, ctypes.c_uint64], i),
        "fjmi_server_poll": ([_vp, _vp, i, _c_int_p], i),
        "fjmi_server_release": ([_vp, ctypes.c_int32], i),
        "fjmi_server_info_get": ([_vp, _vp], i),
        "fjmi_server_destroy": ([_vp], i),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def _check(lib, rc, what):
    if rc != 0:
        msg = lib.fjmi_last_error().decode(errors="replace")
        raise FjmiError(f"{what} failed ({rc}): {msg}")


class Engine:
    """A libfjmi context on one HIP device."""

    def __init__(self, device=0):
        self.lib = load_lib()
        h = _vp()
        _check(self.lib, self.lib.fjmi_create(device, ctypes.byref(h)), "fjmi_create")
        self.ctx = h

    def close(self):
        if self.ctx:
            self.lib.fjmi_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- configuration --
    def kernel_configs(self):
        n = self.lib.fjmi_kernel_configs(None, 0)
        arr = (Bucket * n)()
        self.lib.fjmi_kernel_configs(ctypes.cast(arr, _vp), n)
        return [(b.max_n, b.engine, b.threads, b.slots, b.p4_global) for b in arr]

    def buckets(self):
        arr = (Bucket * 64)()
        n = ctypes.c_int32()
        _check(self.lib, self.lib.fjmi_get_buckets(self.ctx, ctypes.cast(arr, _vp), 64, ctypes.byref(n)),
               "fjmi_get_buckets")
        return [(arr[k].max_n, arr[k].engine, arr[k].threads, arr[k].slots, arr[k].p4_global) for k in range(n.value)]

    def set_buckets(self, rows):
        arr = (Bucket * len(rows))(*[Bucket(*r) for r in rows])
        _check(self.lib, self.lib.fjmi_set_buckets(self.ctx, ctypes.cast(arr, _vp), len(rows)), "fjmi_set_buckets")

    # -- host-memory entry points --
    def cluster_event(self, p4, p, R, engine=ENGINE_AUTO):
        """Clusters one event (N x 4 array of px,py,pz,E). Returns (hist_p1, hist_p2, hist_dij, jets)."""
        p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
        n = len(p4)
        h1 = np.zeros(n, np.int32)
        h2 = np.zeros(n, np.int32)
        hd = np.zeros(n, np.float64)
        jets = np.zeros((max(n, 1), 4), np.float64)
        nj = np.zeros(1, np.int32)
        _check(self.lib, self.lib.fjmi_cluster_event(self.ctx, _ptr(p4), n, int(p), float(R), engine, _ptr(h1),
                                                     _ptr(h2), _ptr(hd), _ptr(jets), _ptr(nj)), "fjmi_cluster_event")
        return h1, h2, hd, jets[:nj[0]].copy()

    def cluster_event_def(self, p4, p, scheme, R, hist_jets=False):
        """Clusters one event under a general jet definition: genkt exponent p (any finite real) and
        recombination scheme SCHEME_*. Returns (hist_p1, hist_p2, hist_dij, jets), and with hist_jets=True
        also the N x 4 momenta each merge step stores (rows of beam steps NaN)."""
        p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
        n = len(p4)
        h1 = np.zeros(n, np.int32)
        h2 = np.zeros(n, np.int32)
        hd = np.zeros(n, np.float64)
        jets = np.zeros((max(n, 1), 4), np.float64)
        nj = np.zeros(1, np.int32)
        hj = np.full((max(n, 1), 4), np.nan) if hist_jets else None
        jd = JetDef(float(p), int(scheme), 0)
        _check(self.lib, self.lib.fjmi_cluster_event_def(self.ctx, _ptr(p4), n, ctypes.byref(jd), float(R), _ptr(h1),
                                                         _ptr(h2), _ptr(hd), _ptr(jets), _ptr(nj), _ptr(hj)),
               "fjmi_cluster_event_def")
        out = (h1, h2, hd, jets[:nj[0]].copy())
        return out + (hj[:n],) if hist_jets else out

    def cluster_batch(self, p4, offsets, p, R):
        """Clusters many events in one graph launch. p and R are scalars or per-event arrays.
        Returns (hist_p1, hist_p2, hist_dij, jets, njets) with per-particle arrays."""
        p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        ne = len(offsets) - 1
        tot = len(p4)
        h1 = np.zeros(tot, np.int32)
        h2 = np.zeros(tot, np.int32)
        hd = np.zeros(tot, np.float64)
        jets = np.zeros((max(tot, 1), 4), np.float64)
        nj = np.zeros(max(ne, 1), np.int32)
        pa = np.ascontiguousarray(p, dtype=np.int32) if np.ndim(p) else None
        Ra = np.ascontiguousarray(R, dtype=np.float64) if np.ndim(R) else None
        _check(self.lib, self.lib.fjmi_cluster_batch(
            self.ctx, ne, _ptr(offsets), _ptr(p4), _ptr(pa), _ptr(Ra), int(p) if pa is None else 0,
            float(R) if Ra is None else 0.0, _ptr(h1), _ptr(h2), _ptr(hd), _ptr(jets), _ptr(nj)),
            "fjmi_cluster_batch")
        return h1, h2, hd, jets, nj[:ne]

    def cluster_batch_def(self, p4, offsets, p, scheme, R, hist_jets=False):
        """Clusters many events from host memory under general jet definitions: p and scheme scalars
        (one definition) or per-event arrays; R scalar or per event.  Returns (hist_p1, hist_p2, hist_dij,
        jets, njets) and, with hist_jets=True, the per-particle history momenta."""
        p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        ne, tot = len(offsets) - 1, len(p4)
        h1 = np.zeros(tot, np.int32)
        h2 = np.zeros(tot, np.int32)
        hd = np.zeros(tot, np.float64)
        jets = np.zeros((max(tot, 1), 4), np.float64)
        nj = np.zeros(max(ne, 1), np.int32)
        hj = np.full((max(tot, 1), 4), np.nan) if hist_jets else None
        defs = None
        jd = JetDef(0.0, 0, 0)
        if np.ndim(p) or np.ndim(scheme):
            defs = np.zeros(ne, JETDEF_DTYPE)
            defs["p"], defs["scheme"] = np.broadcast_to(p, ne), np.broadcast_to(scheme, ne)
        else:
            jd = JetDef(float(p), int(scheme), 0)
        Ra = np.ascontiguousarray(R, dtype=np.float64) if np.ndim(R) else None
        _check(self.lib, self.lib.fjmi_cluster_batch_def(
            self.ctx, ne, _ptr(offsets), _ptr(p4), _ptr(defs), ctypes.byref(jd), _ptr(Ra),
            float(R) if Ra is None else 0.0, _ptr(h1), _ptr(h2), _ptr(hd), _ptr(jets), _ptr(nj), _ptr(hj)),
            "fjmi_cluster_batch_def")
        out = (h1, h2, hd, jets, nj[:ne])
        return out + (hj[:tot],) if hist_jets else out

    # -- device-memory entry points (pointers as ints, e.g. torch tensor data_ptr()) --
    def cluster_event_device(self, d_p4, n, p, R, d_h1, d_h2, d_hd, d_jets, d_nj, stream=0, engine=ENGINE_AUTO):
        _check(self.lib, self.lib.fjmi_cluster_event_device(self.ctx, d_p4, n, int(p), float(R), engine, d_h1, d_h2,
                                                            d_hd, d_jets, d_nj, stream or None),
               "fjmi_cluster_event_device")

    def cluster_event_device_def(self, d_p4, n, p, scheme, R, d_h1, d_h2, d_hd, d_jets, d_nj, d_hist_jets=0,
                                 stream=0):
        jd = JetDef(float(p), int(scheme), 0)
        _check(self.lib, self.lib.fjmi_cluster_event_device_def(self.ctx, d_p4, n, ctypes.byref(jd), float(R), d_h1,
                                                                d_h2, d_hd, d_jets, d_nj, d_hist_jets or None,
                                                                stream or None),
               "fjmi_cluster_event_device_def")

    def cluster_pseudojets(self, d_particles, n, p, R, engine=ENGINE_AUTO, stream=0):
        _check(self.lib, self.lib.fjmi_cluster_pseudojets(self.ctx, d_particles, n, int(p), float(R), engine,
                                                          stream or None), "fjmi_cluster_pseudojets")

    def cluster_event_device_noalloc(self, d_p4, n, p, R, d_h1, d_h2, d_hd, d_jets, d_nj, stream=0):
        """fjmi_cluster_event_device_noalloc: the tiled engine with the library's static scratch (no
        allocation; events of up to 20000 particles; calls must not overlap on one device)."""
        _check(self.lib, self.lib.fjmi_cluster_event_device_noalloc(self.ctx, d_p4, n, int(p), float(R), d_h1, d_h2,
                                                                    d_hd, d_jets, d_nj, stream or None),
               "fjmi_cluster_event_device_noalloc")

    # -- results derived on the device (device pointers as ints; d_offsets 0 for one event) --
    def inclusive_jets(self, n_events, d_offsets, d_jets, d_njets, ptmin, d_out, d_out_n, stream=0):
        _check(self.lib, self.lib.fjmi_inclusive_jets(self.ctx, n_events, d_offsets or None, d_jets, d_njets,
                                                      float(ptmin), d_out, d_out_n, stream or None),
               "fjmi_inclusive_jets")

    def inclusive_pseudojets(self, d_particles, n, ptmin, d_out, d_count, stream=0):
        _check(self.lib, self.lib.fjmi_inclusive_pseudojets(self.ctx, d_particles, n, float(ptmin), d_out, d_count,
                                                            stream or None), "fjmi_inclusive_pseudojets")

    def history_view(self, n_events, d_offsets, n_single, d_h1, d_h2, d_hd, d_child, d_max_dij, stream=0):
        _check(self.lib, self.lib.fjmi_history_view(self.ctx, n_events, d_offsets or None, n_single, d_h1, d_h2,
                                                    d_hd, d_child, d_max_dij, stream or None), "fjmi_history_view")

    def exclusive_jets(self, n_events, d_offsets, n_single, d_p4, d_h1, d_h2, d_hd, mode, param, d_work, d_out,
                       d_out_n, stream=0):
        _check(self.lib, self.lib.fjmi_exclusive_jets(self.ctx, n_events, d_offsets or None, n_single, d_p4, d_h1,
                                                      d_h2, d_hd, int(mode), float(param), d_work, d_out, d_out_n,
                                                      stream or None), "fjmi_exclusive_jets")

    def exclusive_jets_def(self, n_events, d_offsets, n_single, d_p4, d_h1, d_h2, d_hd, d_hist_jets, scheme, mode,
                           param, d_out, d_out_n, stream=0):
        _check(self.lib, self.lib.fjmi_exclusive_jets_def(self.ctx, n_events, d_offsets or None, n_single, d_p4, d_h1,
                                                          d_h2, d_hd, d_hist_jets, int(scheme), int(mode),
                                                          float(param), d_out, d_out_n, stream or None),
               "fjmi_exclusive_jets_def")

    def stream_check(self, stream=0):
        """Synchronises `stream`; raises if a one-event device call on it reported a failure."""
        _check(self.lib, self.lib.fjmi_stream_check(self.ctx, stream or None), "fjmi_stream_check")

    def server(self, **cfg):
        """A continuous-batching Server on this device (see Server)."""
        return Server(self, **cfg)

    def batch_create(self, h_offsets, d_p4, d_offsets, d_p, d_R, p_uni, R_uni, d_h1, d_h2, d_hd, d_jets, d_nj):
        h_offsets = np.ascontiguousarray(h_offsets, dtype=np.int64)
        b = _vp()
        _check(self.lib, self.lib.fjmi_batch_create(self.ctx, len(h_offsets) - 1, _ptr(h_offsets), d_p4, d_offsets,
                                                    d_p, d_R, int(p_uni), float(R_uni), d_h1, d_h2, d_hd, d_jets,
                                                    d_nj, ctypes.byref(b)), "fjmi_batch_create")
        return Batch(self, b)


    def batch_create_def(self, h_offsets, d_p4, d_offsets, d_defs, p, scheme, d_R, R_uni, d_h1, d_h2, d_hd, d_jets,
                         d_nj, d_hist_jets=0):
        """A batch plan under general jet definitions: per event (d_defs, device records of JETDEF_DTYPE)
        or (p, scheme) for all when d_defs is 0; d_hist_jets (optional) gets each merge step's momentum."""
        h_offsets = np.ascontiguousarray(h_offsets, dtype=np.int64)
        b = _vp()
        jd = JetDef(float(p), int(scheme), 0)
        _check(self.lib, self.lib.fjmi_batch_create_def(self.ctx, len(h_offsets) - 1, _ptr(h_offsets), d_p4,
                                                        d_offsets, d_defs or None, ctypes.byref(jd), d_R or None,
                                                        float(R_uni), d_h1, d_h2, d_hd, d_jets, d_nj,
                                                        d_hist_jets or None, ctypes.byref(b)),
               "fjmi_batch_create_def")
        return Batch(self, b)


class Batch:
    def __init__(self, eng, handle):
        self.eng = eng
        self.h = handle

    def launch(self, stream=0):
        _check(self.eng.lib, self.eng.lib.fjmi_batch_launch(self.h, stream or None), "fjmi_batch_launch")

    def sync(self, stream=0):
        _check(self.eng.lib, self.eng.lib.fjmi_batch_sync(self.h, stream or None), "fjmi_batch_sync")

    def stats(self):
        """Per-bucket dicts: max_n, threads, slots, p4_global, n_events, grid, particles, sum_n2, ms."""
        arr = (BucketStats * 64)()
        n = ctypes.c_int32()
        _check(self.eng.lib, self.eng.lib.fjmi_batch_stats(self.h, ctypes.cast(arr, _vp), 64, ctypes.byref(n)),
               "fjmi_batch_stats")
        out = []
        for k in range(n.value):
            b = arr[k]
            out.append(dict(max_n=b.bucket.max_n, engine=b.bucket.engine, threads=b.bucket.threads, slots=b.bucket.slots,
                            p4_global=b.bucket.p4_global, n_events=b.n_events, grid=b.grid,
                            particles=b.particles, sum_n2=b.sum_n2, ms=b.ms))
        return out

    def close(self):
        if self.h:
            self.eng.lib.fjmi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching server on one device (fjmi_server_*, include/fjmi.h).

    submit() admits an event or returns False when the memory budget or the
    event slots are full; poll() returns the completed events (results copied
    into numpy arrays, then released)."""

    def __init__(self, eng, device_budget=0, host_budget=0, max_events=0, variant=1, tile_q=0, idle_us=0):
        self.eng = eng
        cfg = ServerConfig(device_budget, host_budget, max_events, variant, tile_q, idle_us)
        h = _vp()
        _check(eng.lib, eng.lib.fjmi_server_create(eng.ctx, ctypes.byref(cfg), ctypes.byref(h)), "fjmi_server_create")
        self.h = h
        self._res = (ServerResult * 1024)()

    def submit(self, p4, p, R, tag):
        p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
        rc = self.eng.lib.fjmi_server_submit(self.h, _ptr(p4), len(p4), int(p), float(R), int(tag))
        if rc == FJMI_EAGAIN:
            return False
        _check(self.eng.lib, rc, "fjmi_server_submit")
        return True

    def poll(self):
        """Completed events as dicts: tag, n, njets, hist_p1, hist_p2, hist_dij, jets (numpy copies)."""
        out = []
        while True:
            k = ctypes.c_int32()
            _check(self.eng.lib, self.eng.lib.fjmi_server_poll(self.h, ctypes.cast(self._res, _vp), len(self._res),
                                                               ctypes.byref(k)), "fjmi_server_poll")
            for r in self._res[:k.value]:
                n, nj = r.n, r.njets
                out.append(dict(tag=r.tag, n=n, njets=nj,
                                hist_p1=np.ctypeslib.as_array(r.hist_p1, (n,)).copy() if n else np.zeros(0, np.int32),
                                hist_p2=np.ctypeslib.as_array(r.hist_p2, (n,)).copy() if n else np.zeros(0, np.int32),
                                hist_dij=np.ctypeslib.as_array(r.hist_dij, (n,)).copy() if n else np.zeros(0),
                                jets=np.ctypeslib.as_array(r.jets, (nj, 4)).copy() if nj > 0 else np.zeros((0, 4))))
                _check(self.eng.lib, self.eng.lib.fjmi_server_release(self.h, r.handle), "fjmi_server_release")
            if k.value < len(self._res):
                return out

    def info(self):
        i = ServerInfo()
        _check(self.eng.lib, self.eng.lib.fjmi_server_info_get(self.h, ctypes.byref(i)), "fjmi_server_info_get")
        return {f: getattr(i, f) for f, _ in ServerInfo._fields_}

    def close(self):
        if self.h:
            rc = self.eng.lib.fjmi_server_destroy(self.h)
            self.h = None
            _check(self.eng.lib, rc, "fjmi_server_destroy")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- binary event files (fjmi_events_*, include/fjmi.h) --
def write_events(path, p4, offsets, p=None, R=None):
    """Writes a batch (packed (P, 4) four-momenta, int64 offsets[E+1], optional per-event p and R)."""
    lib = load_lib()
    p4 = np.ascontiguousarray(p4, dtype=np.float64).reshape(-1, 4)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pa = None if p is None else np.ascontiguousarray(p, dtype=np.int32)
    Ra = None if R is None else np.ascontiguousarray(R, dtype=np.float64)
    _check(lib, lib.fjmi_events_write(os.fsencode(path), len(offsets) - 1, _ptr(offsets), _ptr(p4), _ptr(pa), _ptr(Ra)),
           "fjmi_events_write")


def read_events(path):
    """(p4 (P, 4), offsets[E+1], p or None, R or None) of an event file."""
    lib = load_lib()
    ne, npart, hp, hr = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
    _check(lib, lib.fjmi_events_info(os.fsencode(path), ctypes.byref(ne), ctypes.byref(npart), ctypes.byref(hp),
                                     ctypes.byref(hr)), "fjmi_events_info")
    offsets = np.zeros(ne.value + 1, np.int64)
    p4 = np.zeros((max(npart.value, 1), 4), np.float64)
    pa = np.zeros(max(ne.value, 1), np.int32) if hp.value else None
    Ra = np.zeros(max(ne.value, 1), np.float64) if hr.value else None
    _check(lib, lib.fjmi_events_read(os.fsencode(path), _ptr(offsets), _ptr(p4), _ptr(pa), _ptr(Ra)),
           "fjmi_events_read")
    return (p4[:npart.value], offsets, None if pa is None else pa[:ne.value], None if Ra is None else Ra[:ne.value])


def read_text_events(path):
    """Events of a text file in the reference driver's format (main.cc:34-70): "px py pz E" per line (extra
    fields ignored); a line that is empty or starts with '#' ends an event; any other line that does not parse
    is an error, as it is there.  Returns (p4 (P, 4), offsets[E+1])."""
    events, cur = [], []
    with open(path) as fh:
        for line in fh:
            fields = line.split()
            try:
                vals = [float(x) for x in fields[:4]] if len(fields) >= 4 else None
            except ValueError:
                vals = None
            if vals is None:
                if line.rstrip("\r\n") and not line.startswith("#"):
                    raise ValueError(f"Error while parsing particles:\n{line.rstrip()}")
                if cur:
                    events.append(cur)
                    cur = []
                continue
            cur.append(vals)
    if cur:
        events.append(cur)
    sizes = np.array([len(e) for e in events], np.int64)
    offsets = np.zeros(len(events) + 1, np.int64)
    np.cumsum(sizes, out=offsets[1:])
    p4 = np.array([v for e in events for v in e], np.float64).reshape(-1, 4)
    return p4, offsets


# -- generator --
def gen_event(workload, event):
    """(N, 4) float64 array of event `event` of workload 1..5 (see csrc/fjmi_gen.c)."""
    g = load_lib("libfjmi_gen.so")
    n = g.fjmi_gen_event_size(workload, event)
    if n < 0:
        raise FjmiError(f"unknown workload {workload}")
    out = np.zeros((max(n, 1), 4), np.float64)
    rc = g.fjmi_gen_event(workload, event, out.ctypes.data_as(_vp), n)
    if rc != n:
        raise FjmiError(f"generator failed ({rc})")
    return out[:n]


def gen_serving_attrs(event):
    """(p, R, gap) of serving request `event`: algorithm and radius drawn from
    {anti-kt 0.4, anti-kt 0.8, C/A 0.8, kt 0.4}, and a unit-mean exponential gap."""
    g = load_lib("libfjmi_gen.so")
    p = ctypes.c_int32()
    R = ctypes.c_double()
    gap = ctypes.c_double()
    g.fjmi_gen_serving_attrs(event, ctypes.byref(p), ctypes.byref(R), ctypes.byref(gap))
    return p.value, R.value, gap.value


def gen_workload(workload, n_events, first=0):
    """Packed (total, 4) array and int64 offsets[n_events+1] for events first..first+n_events-1."""
    g = load_lib("libfjmi_gen.so")
    sizes = np.array([g.fjmi_gen_event_size(workload, first + e) for e in range(n_events)], np.int64)
    offsets = np.zeros(n_events + 1, np.int64)
    np.cumsum(sizes, out=offsets[1:])
    p4 = np.zeros((max(int(offsets[-1]), 1), 4), np.float64)
    for e in range(n_events):
        rc = g.fjmi_gen_event(workload, first + e, p4[offsets[e]:].ctypes.data_as(_vp), int(sizes[e]))
        if rc != sizes[e]:
            raise FjmiError(f"generator failed ({rc})")
    return p4[:offsets[-1]], offsets
