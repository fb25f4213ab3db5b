"""Python host side of the MI355X VELO path.

A thin ctypes layer over ``libvelo_amd.so`` (C ABI declared in ``include/velo_amd.h``).
It mirrors the host side of Allen's VELO-SP sequence
(configuration/python/AllenConf/velo_reconstruction.py:72-321): one call per algorithm,
named like the reference algorithm and taking the reference's argument names, plus
``VeloSequence`` which enqueues the whole path with no host synchronisation.

PyTorch only provides device memory and the HIP stream; every kernel is the library's.
Loading fails loudly when the native library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvelo_amd.so")

VELO_OK = 0
STATUS_TTF_RING_OVERFLOW = 1 << 0
STATUS_TRACK_SLOTS_OVERFLOW = 1 << 1
STATUS_HIT_INDEX_OVERFLOW = 1 << 2
STATUS_THREE_HIT_SLOTS_OVERFLOW = 1 << 3
STATUS_LDS_FALLBACK = 1 << 8
STATUS_OVERFLOW_MASK = 0xFF

P = ctypes.c_void_p
U = ctypes.c_uint
I = ctypes.c_int
F = ctypes.c_float


class VeloError(RuntimeError):
    pass


def _params(name, fields):
    return type(name, (ctypes.Structure,), {"_fields_": fields})


CandidatesParams = _params("CandidatesParams", [
    ("host_number_of_events", U), ("host_raw_bank_version", I), ("dev_event_list", P), ("event_list_size", U),
    ("dev_velo_raw_input", P), ("dev_velo_raw_input_offsets", P), ("dev_velo_raw_input_sizes", P),
    ("dev_velo_raw_input_types", P), ("dev_number_of_candidates", P), ("dev_velo_bank_index", P),
    ("block_dim_x", U), ("event_start", U), ("mep_layout", I)])

EstimateParams = _params("EstimateParams", [
    ("host_number_of_events", U), ("host_number_of_cluster_candidates", U), ("host_raw_bank_version", I),
    ("dev_event_list", P), ("event_list_size", U), ("dev_candidates_offsets", P), ("dev_velo_raw_input", P),
    ("dev_velo_raw_input_offsets", P), ("dev_velo_raw_input_sizes", P), ("dev_velo_raw_input_types", P),
    ("dev_estimated_input_size", P), ("dev_module_candidate_num", P), ("dev_cluster_candidates", P),
    ("event_start", U), ("mep_layout", I)])

MaskedClusteringParams = _params("MaskedClusteringParams", [
    ("host_total_number_of_velo_clusters", U), ("host_number_of_events", U), ("host_raw_bank_version", I),
    ("dev_velo_raw_input", P), ("dev_velo_raw_input_offsets", P), ("dev_velo_raw_input_sizes", P),
    ("dev_velo_raw_input_types", P), ("dev_offsets_estimated_input_size", P), ("dev_module_candidate_num", P),
    ("dev_cluster_candidates", P), ("dev_event_list", P), ("event_list_size", U), ("dev_candidates_offsets", P),
    ("dev_number_of_events", P), ("dev_velo_bank_index", P), ("dev_module_cluster_num", P),
    ("dev_velo_cluster_container", P), ("constants", P), ("event_start", U), ("mep_layout", I),
    ("dev_velo_clusters", P)])

PVParams = _params("PVParams", [
    ("host_number_of_events", U), ("dev_event_list", P), ("event_list_size", U), ("dev_number_of_events", P),
    ("dev_offsets_all_velo_tracks", P), ("dev_velo_kalman_beamline_states", P), ("dev_pvtracks", P),
    ("dev_zhisto", P), ("dev_zpeaks", P), ("dev_number_of_zpeaks", P), ("dev_pvtracks_denom", P),
    ("dev_multi_fit_vertices", P), ("dev_number_of_multi_fit_vertices", P), ("dev_multi_final_vertices", P),
    ("dev_number_of_multi_final_vertices", P), ("dev_status", P), ("constants", P)])

# PV::Vertex (PV_Definitions.cuh:64-97).
PV_VERTEX_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("chi2", "<f4"), ("ndof", "<i4"),
                            ("cov00", "<f4"), ("cov10", "<f4"), ("cov11", "<f4"), ("cov20", "<f4"), ("cov21", "<f4"),
                            ("cov22", "<f4"), ("nTracks", "<f4")])

RetinaSizesParams = _params("RetinaSizesParams", [
    ("host_number_of_events", U), ("host_raw_bank_version", I), ("dev_event_list", P), ("event_list_size", U),
    ("dev_velo_retina_raw_input", P), ("dev_velo_retina_raw_input_offsets", P),
    ("dev_velo_retina_raw_input_sizes", P), ("dev_velo_retina_raw_input_types", P), ("dev_retina_bank_index", P),
    ("dev_each_sensor_pair_size", P), ("event_start", U), ("mep_layout", I)])

RetinaDecodeParams = _params("RetinaDecodeParams", [
    ("host_total_number_of_velo_clusters", U), ("host_raw_bank_version", I), ("host_number_of_events", U),
    ("dev_velo_retina_raw_input", P), ("dev_velo_retina_raw_input_offsets", P),
    ("dev_velo_retina_raw_input_sizes", P), ("dev_velo_retina_raw_input_types", P),
    ("dev_offsets_each_sensor_pair_size", P), ("dev_retina_bank_index", P), ("dev_event_list", P),
    ("event_list_size", U), ("dev_number_of_events", P), ("dev_module_cluster_num", P),
    ("dev_offsets_module_pair_cluster", P), ("dev_velo_cluster_container", P), ("dev_hit_permutations", P),
    ("dev_hit_sorting_key", P), ("constants", P), ("event_start", U), ("mep_layout", I), ("dev_velo_clusters", P)])

SortParams = _params("SortParams", [
    ("host_number_of_events", U), ("host_total_number_of_velo_clusters", U), ("dev_event_list", P),
    ("event_list_size", U), ("dev_offsets_estimated_input_size", P), ("dev_module_cluster_num", P),
    ("dev_velo_cluster_container", P), ("dev_number_of_events", P), ("dev_sorted_velo_cluster_container", P),
    ("dev_hit_permutation", P)])

SearchParams = _params("SearchParams", [
    ("host_number_of_events", U), ("host_total_number_of_velo_clusters", U), ("dev_event_list", P),
    ("event_list_size", U), ("dev_number_of_events", P), ("dev_sorted_velo_cluster_container", P),
    ("dev_offsets_estimated_input_size", P), ("dev_module_cluster_num", P), ("dev_tracks", P),
    ("dev_tracklets", P), ("dev_tracks_to_follow", P), ("dev_three_hit_tracks", P), ("dev_hit_used", P),
    ("dev_atomics_velo", P), ("dev_rel_indices", P), ("dev_number_of_velo_tracks", P), ("dev_status", P),
    ("constants", P), ("phi_tolerance", F), ("max_scatter", F), ("max_skipped_modules", U), ("block_dim_x", U)])

ThreeHitParams = _params("ThreeHitParams", [
    ("host_number_of_events", U), ("host_total_number_of_velo_clusters", U), ("dev_event_list", P),
    ("event_list_size", U), ("dev_sorted_velo_cluster_container", P), ("dev_offsets_estimated_input_size", P),
    ("dev_three_hit_tracks_input", P), ("dev_atomics_velo", P), ("dev_hit_used", P), ("dev_number_of_events", P),
    ("dev_three_hit_tracks_output", P), ("dev_number_of_three_hit_tracks_output", P), ("max_chi2", F)])

CopyHitNumberParams = _params("CopyHitNumberParams", [
    ("host_number_of_events", U), ("host_number_of_velo_tracks_at_least_four_hits", U),
    ("host_number_of_three_hit_tracks_filtered", U), ("dev_tracks", P), ("dev_offsets_velo_tracks", P),
    ("dev_offsets_estimated_input_size", P), ("dev_offsets_number_of_three_hit_tracks_filtered", P),
    ("dev_velo_track_hit_number", P), ("dev_offsets_all_velo_tracks", P),
    ("host_number_of_reconstructed_velo_tracks", P)])

ConsolidateParams = _params("ConsolidateParams", [
    ("host_accumulated_number_of_hits_in_velo_tracks", U), ("host_number_of_reconstructed_velo_tracks", U),
    ("host_number_of_three_hit_tracks_filtered", U), ("host_number_of_events", U), ("dev_event_list", P),
    ("event_list_size", U), ("dev_offsets_all_velo_tracks", P), ("dev_tracks", P),
    ("dev_offsets_velo_track_hit_number", P), ("dev_sorted_velo_cluster_container", P),
    ("dev_offsets_estimated_input_size", P), ("dev_three_hit_tracks_output", P),
    ("dev_offsets_number_of_three_hit_tracks_filtered", P), ("dev_number_of_events", P),
    ("dev_velo_track_hits", P), ("dev_accepted_velo_tracks", P), ("dev_velo_hits_view", P),
    ("dev_velo_track_view", P), ("dev_velo_tracks_view", P), ("dev_velo_multi_event_tracks_view", P),
    ("dev_imec_velo_tracks", P)])

KalmanParams = _params("KalmanParams", [
    ("host_number_of_reconstructed_velo_tracks", U), ("host_number_of_events", U), ("dev_event_list", P),
    ("event_list_size", U), ("dev_number_of_events", P), ("dev_offsets_all_velo_tracks", P),
    ("dev_offsets_velo_track_hit_number", P), ("dev_velo_track_hits", P), ("dev_is_backward", P),
    ("dev_velo_kalman_beamline_states", P), ("dev_velo_kalman_endvelo_states", P), ("constants", P),
    ("covariance_fp64_fallback", U), ("dev_velo_tracks_view", P), ("dev_velo_kalman_beamline_states_view", P),
    ("dev_velo_kalman_endvelo_states_view", P)])

# Allen's device view objects (velo_amd.h): the byte layouts of Velo::Clusters,
# Allen::Views::Velo::Consolidated::{Hits, Track, Tracks}, MultiEventContainer<Tracks> and
# Allen::Views::Physics::KalmanStates with 64-bit pointers; pads are zero.
CLUSTERS_VIEW_DTYPE = np.dtype([("base_pointer", "<u8"), ("total_number_of_hits", "<u4"), ("offset", "<u4")])
HITS_VIEW_DTYPE = np.dtype([("base_pointer", "<u8"), ("offset", "<u4"), ("size", "<u4"),
                            ("total_number_of_hits", "<u4"), ("pad_", "<u4")])
TRACK_VIEW_DTYPE = np.dtype([("hits", "<u8"), ("track_index", "<u4"), ("track_container_offset", "<u4"),
                             ("offset", "<u4"), ("number_of_hits", "<u4")])
TRACKS_VIEW_DTYPE = np.dtype([("track", "<u8"), ("size", "<u4"), ("offset", "<u4")])
MULTI_EVENT_TRACKS_VIEW_DTYPE = np.dtype([("type_id", "<i4"), ("pad_", "<u4"), ("container", "<u8"),
                                          ("number_of_events", "<u4"), ("pad2_", "<u4")])
KALMAN_STATES_VIEW_DTYPE = np.dtype([("base_pointer", "<u8"), ("offset", "<u4"), ("size", "<u4"),
                                     ("total_number_of_tracks", "<u4"), ("pad_", "<u4")])
TYPE_ID_VELO_TRACKS = 1

PrefixSumParams = _params("PrefixSumParams", [
    ("dev_input_buffer", P), ("input_size", U), ("dev_output_buffer", P), ("dev_scratch", P),
    ("host_total_sum_holder", P)])

SequenceInput = _params("SequenceInput", [
    ("dev_velo_raw_input", P), ("dev_velo_raw_input_offsets", P), ("dev_velo_raw_input_sizes", P),
    ("dev_velo_raw_input_types", P), ("raw_input_bytes", ctypes.c_size_t), ("number_of_events", U),
    ("bank_version", I), ("phi_tolerance", F), ("max_scatter", F), ("max_chi2", F), ("max_skipped_modules", U),
    ("retina", I), ("mep_layout", I)])

# Every function the header declares, with its ctypes signature.
SIGNATURES = {
    "velo_last_hip_error": (I, []),
    "velo_error_string": (ctypes.c_char_p, [I]),
    "velo_wall_clock_khz": (I, [ctypes.POINTER(ctypes.c_uint)]),
    "velo_microbench": (I, [ctypes.c_int, ctypes.POINTER(ctypes.c_double), P]),
    "velo_constants_create": (I, [ctypes.c_char_p, ctypes.c_size_t, P, ctypes.POINTER(P)]),
    "velo_constants_destroy": (I, [P]),
    "velo_constants_dev_geometry": (P, [P]),
    "velo_constants_dev_sp_patterns": (P, [P]),
    "velo_constants_dev_sp_fx": (P, [P]),
    "velo_constants_dev_sp_fy": (P, [P]),
    "velo_constants_dev_beamline": (P, [P]),
    "velo_cache_sp_patterns": (None, [P, P, P, P]),
    "velo_constants_sp_tables": (I, [P, P, P, P, P]),
    "velo_prefix_sum_scratch_size": (ctypes.c_size_t, [U]),
    "velo_prefix_sum": (I, [ctypes.POINTER(PrefixSumParams), P]),
    "velo_prefix_sum_static_max_items": (ctypes.c_size_t, []),
    "velo_prefix_sum_static": (I, [ctypes.POINTER(PrefixSumParams), P]),
    "velo_calculate_number_of_candidates": (I, [ctypes.POINTER(CandidatesParams), P]),
    "velo_estimate_input_size": (I, [ctypes.POINTER(EstimateParams), P]),
    "velo_masked_clustering": (I, [ctypes.POINTER(MaskedClusteringParams), P]),
    "velo_retina_sensor_slots": (U, [I]),
    "velo_pv_beamline": (I, [ctypes.POINTER(PVParams), P]),
    "velo_mdf_read": (I, [ctypes.c_char_p, U, U, ctypes.POINTER(P)]),
    "velo_mdf_events_info": (I, [P, ctypes.POINTER(U), ctypes.POINTER(I), ctypes.POINTER(ctypes.c_size_t)]),
    "velo_mdf_events_buffer": (I, [P, ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_size_t)]),
    "velo_mdf_events_free": (None, [P]),
    "velo_calculate_number_of_retinaclusters_each_sensor_pair": (I, [ctypes.POINTER(RetinaSizesParams), P]),
    "velo_decode_retinaclusters": (I, [ctypes.POINTER(RetinaDecodeParams), P]),
    "velo_sort_by_phi": (I, [ctypes.POINTER(SortParams), P]),
    "velo_search_by_triplet": (I, [ctypes.POINTER(SearchParams), P]),
    "velo_track_container_size": (U, [U, U]),
    "velo_three_hit_tracks_filter": (I, [ctypes.POINTER(ThreeHitParams), P]),
    "velo_copy_track_hit_number": (I, [ctypes.POINTER(CopyHitNumberParams), P]),
    "velo_consolidate_tracks": (I, [ctypes.POINTER(ConsolidateParams), P]),
    "velo_kalman_filter": (I, [ctypes.POINTER(KalmanParams), P]),
    "velo_sequence_create": (I, [P, U, ctypes.c_size_t, ctypes.POINTER(P)]),
    "velo_sequence_destroy": (I, [P]),
    "velo_sequence_default_input": (None, [ctypes.POINTER(SequenceInput)]),
    "velo_sequence_run": (I, [P, ctypes.POINTER(SequenceInput), P]),
    "velo_sequence_buffer": (I, [P, ctypes.c_char_p, ctypes.POINTER(P), ctypes.POINTER(ctypes.c_size_t)]),
    "velo_sequence_status": (I, [P, P, ctypes.POINTER(U)]),
    "velo_sequence_set_tuning": (I, [P, ctypes.c_char_p, U]),
    "velo_sequence_stage_ms": (I, [P, ctypes.POINTER(F), ctypes.POINTER(U)]),
    "velo_sequence_capture": (I, [P, ctypes.POINTER(SequenceInput), P]),
    "velo_sequence_replay": (I, [P, P]),
    "velo_sequence_set_scorer": (I, [P, P, F]),
    "velo_mlp_scratch_floats": (ctypes.c_size_t, [U]),
    "velo_mlp_train_step": (I, [P, P, U, P, P, P, P, F, U, P, P, P]),
    "velo_mlp_score": (I, [P, U, P, P, P]),
}

STAGES = {"all": 1, "candidates": 2, "estimate": 3, "clustering": 4, "search": 5, "three_hit_filter": 6,
          "consolidate": 7, "kalman": 8, "pv": 9}

_lib = None


def lib():
    """The native library; raises if it is not built (the product has no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VeloError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = ""):
    if rc != VELO_OK:
        msg = lib().velo_error_string(rc).decode()
        hip = lib().velo_last_hip_error()
        raise VeloError(f"{what}: {msg} (code {rc}, hip error {hip})")


MICROBENCH_KINDS = {"valu_lane_ops_per_s": 0, "lds_bytes_per_s": 1, "hbm_copy_bytes_per_s": 2}


def microbench(kind: str, stream=None) -> float:
    """A measured peak of the current device (velo_microbench): vector-ALU FMA lane operations per
    second, LDS read bytes per second or HBM copy bytes (read + written) per second."""
    out = ctypes.c_double(0.0)
    check(lib().velo_microbench(MICROBENCH_KINDS[kind], ctypes.byref(out), stream), f"microbench {kind}")
    return float(out.value)


def wall_clock_khz() -> int:
    """Tick rate of the device wall clock (kHz), for the "search_event_clock" buffer."""
    khz = ctypes.c_uint(0)
    check(lib().velo_wall_clock_khz(ctypes.byref(khz)), "wall clock rate")
    return int(khz.value)


def cache_sp_patterns():
    """The super-pixel cluster table computed by the library (host only)."""
    pat = np.zeros(256, np.uint8)
    sz = np.zeros(256, np.uint8)
    fx = np.zeros(512, np.float32)
    fy = np.zeros(512, np.float32)
    lib().velo_cache_sp_patterns(pat.ctypes.data, sz.ctypes.data, fx.ctypes.data, fy.ctypes.data)
    return pat, sz, fx, fy


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise VeloError("no HIP device available")
    return torch


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def stream_handle(stream=None) -> int:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


class VeloConstants:
    """Geometry, super-pixel tables and beamline on the device (Allen's Constants for this path)."""

    def __init__(self, geometry: bytes, beamline_xy):
        _torch()
        bl = np.ascontiguousarray(np.asarray(beamline_xy, dtype=np.float32))
        h = P()
        check(lib().velo_constants_create(geometry, len(geometry), bl.ctypes.data, ctypes.byref(h)),
              "velo_constants_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().velo_constants_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sp_tables(self):
        pat = np.zeros(256, np.uint8)
        sz = np.zeros(256, np.uint8)
        fx = np.zeros(512, np.float32)
        fy = np.zeros(512, np.float32)
        check(lib().velo_constants_sp_tables(self.handle, pat.ctypes.data, sz.ctypes.data, fx.ctypes.data,
                                             fy.ctypes.data))
        return pat, sz, fx, fy


class DeviceEvents:
    """Raw banks, offsets, sizes and types of a batch of events resident in HBM."""

    def __init__(self, events, device="cuda"):
        torch = _torch()
        self.n_events = events.n_events
        self.bank_version = events.bank_version
        self.retina = bool(getattr(events, "retina", False))
        self.mep = bool(getattr(events, "mep", False))
        self.raw_bytes = int(events.raw.size)
        self.raw = torch.from_numpy(events.raw).to(device)
        self.offsets = torch.from_numpy(events.offsets.view(np.int32)).to(device)
        self.sizes = torch.from_numpy(events.sizes.view(np.int32)).to(device)
        self.types = torch.from_numpy(events.types.view(np.int32)).to(device)


def prefix_sum(dev_input, n: int, dev_output, stream=None, want_total=False):
    """host_prefix_sum_t as a device scan (dev_output gets n + 1 entries)."""
    torch = _torch()
    scratch_words = lib().velo_prefix_sum_scratch_size(n)
    scratch = torch.empty(max(scratch_words, 1), dtype=torch.int32, device=dev_output.device)
    total = np.zeros(1, np.uint32)
    p = PrefixSumParams(_ptr(dev_input), n, _ptr(dev_output), _ptr(scratch),
                        total.ctypes.data if want_total else 0)
    check(lib().velo_prefix_sum(ctypes.byref(p), stream_handle(stream)), "velo_prefix_sum")
    return int(total[0]) if want_total else None


def prefix_sum_static(dev_input, n: int, dev_output, stream=None, want_total=False):
    """The same device scan with the library's static scratch (allocates nothing)."""
    total = np.zeros(1, np.uint32)
    p = PrefixSumParams(_ptr(dev_input), n, _ptr(dev_output), 0, total.ctypes.data if want_total else 0)
    check(lib().velo_prefix_sum_static(ctypes.byref(p), stream_handle(stream)), "velo_prefix_sum_static")
    return int(total[0]) if want_total else None


class VeloSequence:
    """The whole VELO path enqueued on one stream (no host synchronisation inside a run)."""

    def __init__(self, constants: VeloConstants, max_events: int, max_raw_bytes: int):
        _torch()
        self.constants = constants
        h = P()
        check(lib().velo_sequence_create(constants.handle, max_events, max_raw_bytes, ctypes.byref(h)),
              "velo_sequence_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().velo_sequence_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, knob: str, value: int):
        check(lib().velo_sequence_set_tuning(self.handle, knob.encode(), value), f"tuning {knob}")

    def run(self, ev: DeviceEvents, stream=None, **props):
        inp = SequenceInput()
        lib().velo_sequence_default_input(ctypes.byref(inp))
        inp.dev_velo_raw_input = _ptr(ev.raw)
        inp.dev_velo_raw_input_offsets = _ptr(ev.offsets)
        inp.dev_velo_raw_input_sizes = _ptr(ev.sizes)
        inp.dev_velo_raw_input_types = _ptr(ev.types)
        inp.raw_input_bytes = ev.raw_bytes
        inp.number_of_events = ev.n_events
        inp.bank_version = ev.bank_version
        inp.retina = int(ev.retina)
        inp.mep_layout = int(ev.mep)
        for k, v in props.items():
            setattr(inp, k, v)
        check(lib().velo_sequence_run(self.handle, ctypes.byref(inp), stream_handle(stream)), "velo_sequence_run")

    def capture(self, ev: DeviceEvents, stream, **props):
        """Captures one run into a HIP graph on `stream` (a torch.cuda.Stream, not the null stream)."""
        inp = SequenceInput()
        lib().velo_sequence_default_input(ctypes.byref(inp))
        inp.dev_velo_raw_input = _ptr(ev.raw)
        inp.dev_velo_raw_input_offsets = _ptr(ev.offsets)
        inp.dev_velo_raw_input_sizes = _ptr(ev.sizes)
        inp.dev_velo_raw_input_types = _ptr(ev.types)
        inp.raw_input_bytes = ev.raw_bytes
        inp.number_of_events = ev.n_events
        inp.bank_version = ev.bank_version
        inp.retina = int(ev.retina)
        inp.mep_layout = int(ev.mep)
        for k, v in props.items():
            setattr(inp, k, v)
        check(lib().velo_sequence_capture(self.handle, ctypes.byref(inp), stream_handle(stream)),
              "velo_sequence_capture")

    def seed_log(self):
        """(records u32 [n, 5], features f32 [n, 10], seeds seen) of the last run's seed log."""
        seen = int(self.fetch("seed_log_counter", np.uint32, 1)[0])
        cap = self.buffer_ptr("seed_log_records")[1] // 20
        n = min(seen, cap)
        rec = self.fetch("seed_log_records", np.uint32, 5 * n).reshape(n, 5)
        feat = self.fetch("seed_log_features", np.float32, 10 * n).reshape(n, 10)
        return rec, feat, seen

    def set_scorer(self, params, threshold: float = 0.0):
        """NN-pruned seeding with these scorer parameters (None: plain search)."""
        if params is None:
            check(lib().velo_sequence_set_scorer(self.handle, None, 0.0), "set_scorer")
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        check(lib().velo_sequence_set_scorer(self.handle, p.ctypes.data, float(threshold)), "set_scorer")

    def seed_logits(self, n: int) -> np.ndarray:
        return self.fetch("seed_log_logits", np.float32, n)

    def replay(self, stream):
        """Launches the captured run on `stream`."""
        check(lib().velo_sequence_replay(self.handle, stream_handle(stream)), "velo_sequence_replay")

    def time_stage(self, stage: str | None):
        """Brackets one stage of subsequent runs with HIP events (None: no timing)."""
        self.set_tuning("timed_stage", STAGES[stage] if stage else 0)

    def stage_ms(self):
        """(total milliseconds, runs) of the timed stage since time_stage() was called."""
        ms = F()
        runs = U()
        check(lib().velo_sequence_stage_ms(self.handle, ctypes.byref(ms), ctypes.byref(runs)), "stage_ms")
        return ms.value, runs.value

    def buffer_ptr(self, name: str):
        p = P()
        n = ctypes.c_size_t()
        check(lib().velo_sequence_buffer(self.handle, name.encode(), ctypes.byref(p), ctypes.byref(n)), name)
        return p.value, n.value

    def status(self, stream=None) -> int:
        st = U()
        check(lib().velo_sequence_status(self.handle, stream_handle(stream), ctypes.byref(st)), "status")
        return st.value

    def fetch(self, name: str, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        """Copies `count` elements of a sequence buffer to the host (synchronising)."""
        torch = _torch()
        ptr, cap = self.buffer_ptr(name)
        dt = np.dtype(dtype)
        if count is None:
            count = cap // dt.itemsize
        nbytes = count * dt.itemsize
        if offset * dt.itemsize + nbytes > cap:
            raise VeloError(f"{name}: {count} elements exceed the capacity")
        out = np.empty(count, dtype=dt)
        if nbytes:
            torch.cuda.synchronize()
            _hip_memcpy_d2h(out.ctypes.data, ptr + offset * dt.itemsize, nbytes)
        return out


_hip = None


def _hip_memcpy_d2h(dst: int, src: int, nbytes: int):
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        _hip.hipMemcpy.restype = ctypes.c_int
        _hip.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
    rc = _hip.hipMemcpy(dst, src, nbytes, 2)  # hipMemcpyDeviceToHost
    if rc != 0:
        raise VeloError(f"hipMemcpy failed ({rc})")


TRACK_DTYPE = np.dtype([("hits", "<u2", (26,)), ("hitsNum", "u1"), ("pad", "u1")])
TRACKLET_DTYPE = np.dtype([("hits", "<u2", (3,))])
