"""TEST INFRASTRUCTURE — Python access to the CPU oracle of the VELO path.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import this
package.  It loads ``oracle/build/libvelo_oracle.so`` (built by ``oracle/Makefile``), a
from-scratch C++ restatement of Allen's VELO-SP sequence with CPU-target semantics.
What pins each part, and which parts are "parity unpinned", is listed in the header of
``oracle/common.hpp``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# VELO_ORACLE_LIB selects another build of the oracle (the planted-mistake variants of
# tests/test_pins_mutants.py); the default is the oracle itself.
LIB_PATH = os.environ.get("VELO_ORACLE_LIB") or os.path.join(_HERE, "build", "libvelo_oracle.so")


class VoInputs(ctypes.Structure):
    _fields_ = [
        ("raw", ctypes.c_void_p),
        ("offsets", ctypes.c_void_p),
        ("sizes", ctypes.c_void_p),
        ("types", ctypes.c_void_p),
        ("n_events", ctypes.c_uint),
        ("event_start", ctypes.c_uint),
        ("bank_version", ctypes.c_int),
        ("geometry", ctypes.c_char_p),
        ("geometry_bytes", ctypes.c_size_t),
        ("beamline", ctypes.c_void_p),
        ("phi_tolerance", ctypes.c_float),
        ("max_scatter", ctypes.c_float),
        ("max_skipped_modules", ctypes.c_uint),
        ("max_chi2", ctypes.c_float),
        ("block_dim_x", ctypes.c_uint),
        ("compute_f64", ctypes.c_int),
        ("kalman_fp64_fallback", ctypes.c_int),
        ("collect_seeds", ctypes.c_int),
        ("retina", ctypes.c_int),
        ("pv_finder", ctypes.c_int),
        ("event_list", ctypes.c_void_p),
        ("event_list_size", ctypes.c_uint),
        ("vectorized_seeding", ctypes.c_int),
        ("sorted_clusters", ctypes.c_void_p),
        ("clusters_offsets", ctypes.c_void_p),
        ("clusters_counts", ctypes.c_void_p),
        ("seed_logits", ctypes.c_void_p),
        ("seed_threshold", ctypes.c_float),
        ("seed_block", ctypes.c_uint),
        ("mep_layout", ctypes.c_int),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.vo_default_inputs.argtypes = [ctypes.POINTER(VoInputs)]
        L.vo_run.restype = ctypes.c_void_p
        L.vo_run.argtypes = [ctypes.POINTER(VoInputs)]
        L.vo_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                             ctypes.POINTER(ctypes.c_size_t)]
        L.vo_free.argtypes = [ctypes.c_void_p]
        L.vo_throughput.restype = ctypes.c_double
        L.vo_throughput.argtypes = [ctypes.POINTER(VoInputs), ctypes.c_uint, ctypes.c_uint,
                                    ctypes.POINTER(ctypes.c_double)]
        L.vo_cache_sp_patterns.argtypes = [ctypes.c_void_p] * 4
        L.vo_fast_atan2f.restype = ctypes.c_float
        L.vo_fast_atan2f.argtypes = [ctypes.c_float, ctypes.c_float]
        L.vo_hit_phi_16.restype = ctypes.c_int16
        L.vo_hit_phi_16.argtypes = [ctypes.c_float, ctypes.c_float]
        L.vo_hit_phi_float_to_16.restype = ctypes.c_int16
        L.vo_hit_phi_float_to_16.argtypes = [ctypes.c_float]
        L.vo_prefix_sum.restype = ctypes.c_uint32
        L.vo_prefix_sum.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
        L.vo_binary_search_leftmost_i16.restype = ctypes.c_int
        L.vo_binary_search_leftmost_i16.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_int16]
        L.vo_make_cluster.restype = ctypes.c_uint64
        L.vo_make_cluster.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint,
                                      ctypes.c_void_p]
        L.vo_three_hit_chi2.restype = ctypes.c_float
        L.vo_three_hit_chi2.argtypes = [ctypes.c_void_p]
        L.vo_load_geometry.restype = ctypes.c_int
        L.vo_load_geometry.argtypes = [ctypes.c_char_p, ctypes.c_size_t] + [ctypes.c_void_p] * 4
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class Result:
    """Every buffer of one oracle run, fetched by the reference's argument name."""

    DTYPES = {
        "tracks": np.dtype([("hits", "<u2", (26,)), ("hitsNum", "u1"), ("pad", "u1")]),
        "three_hit_tracks": np.dtype([("hits", "<u2", (3,))]),
        "three_hit_tracks_output": np.dtype([("hits", "<u2", (3,))]),
        "tracklets": np.dtype([("hits", "<u2", (3,))]),
        "hit_used": np.bool_, "is_backward": np.bool_, "rel_indices": np.uint16,
        "velo_cluster_container": np.uint8, "sorted_velo_cluster_container": np.uint8,
        "velo_track_hits": np.uint8,
        "velo_kalman_beamline_states": np.float32, "velo_kalman_endvelo_states": np.float32,
        "velo_kalman_beamline_states_f64": np.float64, "velo_kalman_endvelo_states_f64": np.float64,
        "search_stats": np.uint64,
        "seed_features": np.float32,
        "hit_sorting_key": np.int64,
        "pvtracks": np.float32, "zhisto": np.float32, "zpeaks": np.float32, "pvtracks_denom": np.float32,
        "multi_fit_vertices": np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("chi2", "<f4"), ("ndof", "<i4"),
                                        ("cov00", "<f4"), ("cov10", "<f4"), ("cov11", "<f4"), ("cov20", "<f4"),
                                        ("cov21", "<f4"), ("cov22", "<f4"), ("nTracks", "<f4")]),
    }
    DTYPES["multi_final_vertices"] = DTYPES["multi_fit_vertices"]

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep  # input arrays must outlive the handle

    def __getitem__(self, name) -> np.ndarray:
        p = ctypes.c_void_p()
        n = ctypes.c_size_t()
        if lib().vo_get(self._h, name.encode(), ctypes.byref(p), ctypes.byref(n)):
            raise KeyError(name)
        dt = np.dtype(self.DTYPES.get(name, np.uint32))
        if n.value == 0:
            return np.zeros(0, dtype=dt)
        buf = (ctypes.c_char * n.value).from_address(p.value)
        return np.frombuffer(bytes(buf), dtype=dt).copy()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().vo_free(self._h)
            self._h = None


def make_inputs(events, geometry: bytes, beamline: np.ndarray, **props):
    a = VoInputs()
    lib().vo_default_inputs(ctypes.byref(a))
    keep = [events.raw, events.offsets, events.sizes, events.types, beamline, geometry]
    a.raw = _ptr(events.raw)
    a.offsets = _ptr(events.offsets)
    a.sizes = _ptr(events.sizes)
    a.types = _ptr(events.types)
    a.n_events = events.n_events
    a.bank_version = events.bank_version
    a.geometry = geometry
    a.geometry_bytes = len(geometry)
    a.beamline = _ptr(beamline)
    a.retina = int(getattr(events, "retina", False))
    a.mep_layout = int(getattr(events, "mep", False))
    event_list = props.pop("event_list", None)
    if event_list is not None:
        event_list = np.ascontiguousarray(event_list, dtype=np.uint32)
        keep.append(event_list)
        a.event_list = _ptr(event_list)
        a.event_list_size = event_list.size
    seed_logits = props.pop("seed_logits", None)
    if seed_logits is not None:
        seed_logits = np.ascontiguousarray(seed_logits, dtype=np.float32)
        keep.append(seed_logits)
        a.seed_logits = _ptr(seed_logits)
    for k, v in props.items():
        setattr(a, k, v)
    return a, keep


def run_clusters(sorted_container, offsets, counts, n_events, geometry: bytes, beamline: np.ndarray, **props):
    """The sequence from velo_search_by_triplet on, over pre-built sorted clusters (Velo::Clusters
    layout, uint8) with their module-pair offsets (26 n + 1) and counts (26 n)."""
    a = VoInputs()
    lib().vo_default_inputs(ctypes.byref(a))
    sorted_container = np.ascontiguousarray(sorted_container, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint32)
    counts = np.ascontiguousarray(counts, dtype=np.uint32)
    assert offsets.size == 26 * n_events + 1 and counts.size == 26 * n_events
    assert sorted_container.size >= 18 * int(offsets[-1])
    keep = [sorted_container, offsets, counts, beamline, geometry]
    a.sorted_clusters, a.clusters_offsets, a.clusters_counts = _ptr(sorted_container), _ptr(offsets), _ptr(counts)
    a.n_events = n_events
    a.geometry = geometry
    a.geometry_bytes = len(geometry)
    a.beamline = _ptr(beamline)
    for k, v in props.items():
        setattr(a, k, v)
    h = lib().vo_run(ctypes.byref(a))
    if not h:
        raise ValueError("oracle run failed")
    return Result(h, keep)


def run(events, geometry: bytes, beamline: np.ndarray, **props) -> Result:
    a, keep = make_inputs(events, geometry, beamline, **props)
    h = lib().vo_run(ctypes.byref(a))
    if not h:
        raise ValueError("oracle run failed (bank version or geometry)")
    return Result(h, keep)


NATIVE_LIB_PATH = os.path.join(_HERE, "build", "libvelo_oracle_native.so")
_native = None


def native_lib():
    """The -O3 -march=x86-64-v3 build of the same sources (oracle/Makefile `native`), used only to
    time the CPU baseline."""
    global _native
    if _native is None:
        if not os.path.exists(NATIVE_LIB_PATH):
            raise RuntimeError(f"{NATIVE_LIB_PATH} is not built; run `make -C oracle native`")
        L = ctypes.CDLL(NATIVE_LIB_PATH)
        L.vo_default_inputs.argtypes = [ctypes.POINTER(VoInputs)]
        L.vo_throughput.restype = ctypes.c_double
        L.vo_throughput.argtypes = [ctypes.POINTER(VoInputs), ctypes.c_uint, ctypes.c_uint,
                                    ctypes.POINTER(ctypes.c_double)]
        _native = L
    return _native


def throughput(events, geometry: bytes, beamline: np.ndarray, n_threads: int, repetitions: int = 1,
               native: bool = False, **props):
    """Events per second of the sequence over the events, one event per task on n_threads
    std::threads, `repetitions` passes; native selects the -O3 x86-64-v3 build."""
    a, keep = make_inputs(events, geometry, beamline, **props)
    secs = ctypes.c_double()
    L = native_lib() if native else lib()
    eps = L.vo_throughput(ctypes.byref(a), n_threads, repetitions, ctypes.byref(secs))
    del keep
    return eps, secs.value


def sp_tables():
    pat = np.zeros(256, np.uint8)
    sz = np.zeros(256, np.uint8)
    fx = np.zeros(512, np.float32)
    fy = np.zeros(512, np.float32)
    lib().vo_cache_sp_patterns(_ptr(pat), _ptr(sz), _ptr(fx), _ptr(fy))
    return pat, sz, fx, fy
