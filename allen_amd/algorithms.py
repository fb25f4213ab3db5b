"""Allen's VELO-SP sequence on the algorithm-by-algorithm C ABI.

This is the host side of configuration/python/AllenConf/velo_reconstruction.py:72-321 (decode_velo
with retina_decoding=False, make_velo_tracks, run_velo_kalman_filter) as Allen's scheduler runs it:
each algorithm's outputs are sized on the host from the previous algorithms' host totals (their
``set_arguments_size``), and each ``host_prefix_sum_t`` is a device scan whose total is copied back
(HostPrefixSum.cpp:76-107). Buffers live in an argument store, a dict of device tensors keyed by the
reference's argument names without the ``dev_``/``host_`` prefix.

The whole-sequence entry point (``velo.VeloSequence``) is what runs without host synchronisation;
this module is the drop-in path an Allen sequence binds, one call per algorithm, with the event
list (MASK_INPUT) and ``event_start`` (RuntimeOptions::event_interval) of the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from allen_amd import velo

N_MODULE_PAIRS = 26
N_SENSORS = 208
MAX_TRACKS_TO_FOLLOW = 2048  # VeloDefinitions.cuh:26
NUM_ATOMICS = 3               # Velo::num_atomics
TRACK_BYTES = 54              # Velo::TrackHits: 26 x u16 + u8 (+ pad)
TRACKLET_BYTES = 6            # Velo::TrackletHits: 3 x u16
CLUSTER_BYTES = 18            # Velo::Clusters element with half_t == float
STATE_FLOATS = 14             # Velo::Consolidated::States: 6 state + 8 covariance words per track


class Arguments(dict):
    """Device buffers by argument name, plus the host scalars the sizes come from."""

    def __init__(self, device="cuda", poison=False):
        super().__init__()
        self.device = device
        self.host = {}
        # poison: outputs start as all-ones bytes instead of whatever the allocator hands out, so
        # that a value an algorithm should write (or clear) and does not shows up.
        self.poison = poison

    def make(self, name, count, dtype=None, fill=None):
        torch = velo._torch()
        dtype = dtype or torch.int32
        if fill is None and self.poison:
            t = torch.full((max(int(count), 1),), -1 if dtype != torch.uint8 else 255, dtype=dtype,
                           device=self.device)
        elif fill is None:
            t = torch.empty(max(int(count), 1), dtype=dtype, device=self.device)
        else:
            t = torch.full((max(int(count), 1),), fill, dtype=dtype, device=self.device)
        self[name] = t
        return t

    def ptr(self, name):
        return int(self[name].data_ptr())


def _list_fields(p, args):
    p.dev_event_list = args.ptr("event_list")
    p.event_list_size = int(args.host["event_list_size"])


def host_prefix_sum(args, input_name, output_name, n):
    """host_prefix_sum_t (HostPrefixSum.cpp:76-107): output of n + 1 entries, total to the host."""
    torch = velo._torch()
    args.make(output_name, n + 1)
    return velo.prefix_sum(args[input_name], n, args[output_name], want_total=True)


def velo_calculate_number_of_candidates_t(args, events, block_dim_x=256):
    """VeloCalculateNumberOfCandidates.cu:50-86."""
    n = args.host["number_of_events"]
    args.make("number_of_candidates", n)
    args.make("velo_bank_index", N_SENSORS, fill=0)
    p = velo.CandidatesParams()
    p.host_number_of_events, p.host_raw_bank_version = n, events.bank_version
    _list_fields(p, args)
    p.dev_velo_raw_input, p.dev_velo_raw_input_offsets = events.raw.data_ptr(), events.offsets.data_ptr()
    p.dev_velo_raw_input_sizes, p.dev_velo_raw_input_types = events.sizes.data_ptr(), events.types.data_ptr()
    p.dev_number_of_candidates, p.dev_velo_bank_index = args.ptr("number_of_candidates"), args.ptr("velo_bank_index")
    p.block_dim_x, p.event_start, p.mep_layout = block_dim_x, args.host["event_start"], int(events.mep)
    velo.check(velo.lib().velo_calculate_number_of_candidates(ctypes.byref(p), None),
               "velo_calculate_number_of_candidates")


def velo_estimate_input_size_t(args, events):
    """EstimateInputSize.cu:241-281."""
    n = args.host["number_of_events"]
    n_cand = args.host["number_of_cluster_candidates"]
    args.make("estimated_input_size", N_MODULE_PAIRS * n)
    args.make("module_candidate_num", n)
    args.make("cluster_candidates", n_cand)
    p = velo.EstimateParams()
    p.host_number_of_events, p.host_number_of_cluster_candidates = n, n_cand
    p.host_raw_bank_version = events.bank_version
    _list_fields(p, args)
    p.dev_candidates_offsets = args.ptr("candidates_offsets")
    p.dev_velo_raw_input, p.dev_velo_raw_input_offsets = events.raw.data_ptr(), events.offsets.data_ptr()
    p.dev_velo_raw_input_sizes, p.dev_velo_raw_input_types = events.sizes.data_ptr(), events.types.data_ptr()
    p.dev_estimated_input_size = args.ptr("estimated_input_size")
    p.dev_module_candidate_num = args.ptr("module_candidate_num")
    p.dev_cluster_candidates = args.ptr("cluster_candidates")
    p.event_start, p.mep_layout = args.host["event_start"], int(events.mep)
    velo.check(velo.lib().velo_estimate_input_size(ctypes.byref(p), None), "velo_estimate_input_size")


def velo_masked_clustering_t(args, events, constants):
    """MaskedVeloClustering.cu:578-619."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    total = args.host["total_number_of_velo_clusters"]
    args.make("module_cluster_num", N_MODULE_PAIRS * n)
    args.make("velo_cluster_container", CLUSTER_BYTES * total, torch.uint8)
    p = velo.MaskedClusteringParams()
    p.host_total_number_of_velo_clusters, p.host_number_of_events = total, n
    p.host_raw_bank_version = events.bank_version
    p.dev_velo_raw_input, p.dev_velo_raw_input_offsets = events.raw.data_ptr(), events.offsets.data_ptr()
    p.dev_velo_raw_input_sizes, p.dev_velo_raw_input_types = events.sizes.data_ptr(), events.types.data_ptr()
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_module_candidate_num = args.ptr("module_candidate_num")
    p.dev_cluster_candidates = args.ptr("cluster_candidates")
    _list_fields(p, args)
    p.dev_candidates_offsets = args.ptr("candidates_offsets")
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_velo_bank_index = args.ptr("velo_bank_index")
    p.dev_module_cluster_num = args.ptr("module_cluster_num")
    p.dev_velo_cluster_container = args.ptr("velo_cluster_container")
    p.constants = constants.handle
    p.event_start, p.mep_layout = args.host["event_start"], int(events.mep)
    args.make("velo_clusters", n * velo.CLUSTERS_VIEW_DTYPE.itemsize, torch.uint8)
    p.dev_velo_clusters = args.ptr("velo_clusters")
    velo.check(velo.lib().velo_masked_clustering(ctypes.byref(p), None), "velo_masked_clustering")


def velo_sort_by_phi_t(args):
    """SortByPhi.cu:14-44."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    total = args.host["total_number_of_velo_clusters"]
    args.make("sorted_velo_cluster_container", CLUSTER_BYTES * total, torch.uint8)
    args.make("hit_permutation", total)
    p = velo.SortParams()
    p.host_number_of_events, p.host_total_number_of_velo_clusters = n, total
    _list_fields(p, args)
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_module_cluster_num = args.ptr("module_cluster_num")
    p.dev_velo_cluster_container = args.ptr("velo_cluster_container")
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_sorted_velo_cluster_container = args.ptr("sorted_velo_cluster_container")
    p.dev_hit_permutation = args.ptr("hit_permutation")
    velo.check(velo.lib().velo_sort_by_phi(ctypes.byref(p), None), "velo_sort_by_phi")


def velo_search_by_triplet_t(args, constants, phi_tolerance=0.045, max_scatter=0.08, max_skipped_modules=1,
                             block_dim_x=0):
    """SearchByTriplet.cu:18-63."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    total = args.host["total_number_of_velo_clusters"]
    tcs = velo.lib().velo_track_container_size(total, n)
    args.make("tracks", tcs * TRACK_BYTES, torch.uint8)
    args.make("tracklets", n * MAX_TRACKS_TO_FOLLOW * TRACKLET_BYTES, torch.uint8)
    args.make("tracks_to_follow", n * MAX_TRACKS_TO_FOLLOW)
    args.make("three_hit_tracks", tcs * TRACKLET_BYTES, torch.uint8)
    args.make("hit_used", total, torch.uint8)
    args.make("atomics_velo", n * NUM_ATOMICS)
    args.make("rel_indices", total, torch.int16)
    args.make("number_of_velo_tracks", n)
    args.make("search_status", 1, fill=0)
    p = velo.SearchParams()
    p.host_number_of_events, p.host_total_number_of_velo_clusters = n, total
    _list_fields(p, args)
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_sorted_velo_cluster_container = args.ptr("sorted_velo_cluster_container")
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_module_cluster_num = args.ptr("module_cluster_num")
    for name in ["tracks", "tracklets", "tracks_to_follow", "three_hit_tracks", "hit_used", "atomics_velo",
                 "rel_indices", "number_of_velo_tracks"]:
        setattr(p, "dev_" + name, args.ptr(name))
    p.dev_status = args.ptr("search_status")
    p.constants = constants.handle
    p.phi_tolerance, p.max_scatter = phi_tolerance, max_scatter
    p.max_skipped_modules, p.block_dim_x = max_skipped_modules, block_dim_x
    velo.check(velo.lib().velo_search_by_triplet(ctypes.byref(p), None), "velo_search_by_triplet")


def velo_three_hit_tracks_filter_t(args, max_chi2=20.0):
    """ThreeHitTracksFilter.cu:14-45."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    total = args.host["total_number_of_velo_clusters"]
    tcs = velo.lib().velo_track_container_size(total, n)
    args.make("three_hit_tracks_output", tcs * TRACKLET_BYTES, torch.uint8)
    args.make("number_of_three_hit_tracks_output", n)
    p = velo.ThreeHitParams()
    p.host_number_of_events, p.host_total_number_of_velo_clusters = n, total
    _list_fields(p, args)
    p.dev_sorted_velo_cluster_container = args.ptr("sorted_velo_cluster_container")
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_three_hit_tracks_input = args.ptr("three_hit_tracks")
    p.dev_atomics_velo, p.dev_hit_used = args.ptr("atomics_velo"), args.ptr("hit_used")
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_three_hit_tracks_output = args.ptr("three_hit_tracks_output")
    p.dev_number_of_three_hit_tracks_output = args.ptr("number_of_three_hit_tracks_output")
    p.max_chi2 = max_chi2
    velo.check(velo.lib().velo_three_hit_tracks_filter(ctypes.byref(p), None), "velo_three_hit_tracks_filter")


def velo_copy_track_hit_number_t(args):
    """VeloCopyTrackHitNumber.cu:10-47; returns host_number_of_reconstructed_velo_tracks."""
    n = args.host["number_of_events"]
    n4 = args.host["number_of_velo_tracks_at_least_four_hits"]
    n3 = args.host["number_of_three_hit_tracks_filtered"]
    args.make("velo_track_hit_number", n4 + n3)
    args.make("offsets_all_velo_tracks", n + 1)
    total = np.zeros(1, np.uint32)
    p = velo.CopyHitNumberParams()
    p.host_number_of_events = n
    p.host_number_of_velo_tracks_at_least_four_hits, p.host_number_of_three_hit_tracks_filtered = n4, n3
    p.dev_tracks, p.dev_offsets_velo_tracks = args.ptr("tracks"), args.ptr("offsets_velo_tracks")
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_offsets_number_of_three_hit_tracks_filtered = args.ptr("offsets_three_hit_tracks_filtered")
    p.dev_velo_track_hit_number = args.ptr("velo_track_hit_number")
    p.dev_offsets_all_velo_tracks = args.ptr("offsets_all_velo_tracks")
    p.host_number_of_reconstructed_velo_tracks = total.ctypes.data
    velo.check(velo.lib().velo_copy_track_hit_number(ctypes.byref(p), None), "velo_copy_track_hit_number")
    return int(total[0])


def velo_consolidate_tracks_t(args):
    """VeloConsolidateTracks.cu:48-110: the track hits, then the view objects of create_velo_views."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    nt = args.host["number_of_reconstructed_velo_tracks"]
    hits = args.host["accumulated_number_of_hits_in_velo_tracks"]
    args.make("velo_track_hits", CLUSTER_BYTES * hits, torch.uint8)
    args.make("accepted_velo_tracks", nt, torch.uint8)
    p = velo.ConsolidateParams()
    p.host_accumulated_number_of_hits_in_velo_tracks, p.host_number_of_reconstructed_velo_tracks = hits, nt
    p.host_number_of_three_hit_tracks_filtered = args.host["number_of_three_hit_tracks_filtered"]
    p.host_number_of_events = n
    _list_fields(p, args)
    p.dev_offsets_all_velo_tracks, p.dev_tracks = args.ptr("offsets_all_velo_tracks"), args.ptr("tracks")
    p.dev_offsets_velo_track_hit_number = args.ptr("offsets_velo_track_hit_number")
    p.dev_sorted_velo_cluster_container = args.ptr("sorted_velo_cluster_container")
    p.dev_offsets_estimated_input_size = args.ptr("offsets_estimated_input_size")
    p.dev_three_hit_tracks_output = args.ptr("three_hit_tracks_output")
    p.dev_offsets_number_of_three_hit_tracks_filtered = args.ptr("offsets_three_hit_tracks_filtered")
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_velo_track_hits, p.dev_accepted_velo_tracks = args.ptr("velo_track_hits"), args.ptr("accepted_velo_tracks")
    args.make("velo_hits_view", n * velo.HITS_VIEW_DTYPE.itemsize, torch.uint8)
    args.make("velo_track_view", nt * velo.TRACK_VIEW_DTYPE.itemsize, torch.uint8)
    args.make("velo_tracks_view", n * velo.TRACKS_VIEW_DTYPE.itemsize, torch.uint8)
    args.make("velo_multi_event_tracks_view", velo.MULTI_EVENT_TRACKS_VIEW_DTYPE.itemsize, torch.uint8)
    args.make("imec_velo_tracks", 8, torch.uint8)
    for name in ["velo_hits_view", "velo_track_view", "velo_tracks_view", "velo_multi_event_tracks_view",
                 "imec_velo_tracks"]:
        setattr(p, "dev_" + name, args.ptr(name))
    velo.check(velo.lib().velo_consolidate_tracks(ctypes.byref(p), None), "velo_consolidate_tracks")


def velo_kalman_filter_t(args, constants, covariance_fp64_fallback=0, from_views=True):
    """VeloKalmanFilter.cu:9-31, 152-228. from_views: the tracks come through dev_velo_tracks_view,
    as in the reference; otherwise through the hit-number offsets and the track hits."""
    torch = velo._torch()
    n = args.host["number_of_events"]
    nt = args.host["number_of_reconstructed_velo_tracks"]
    args.make("is_backward", nt, torch.uint8)
    args.make("velo_kalman_beamline_states", STATE_FLOATS * nt, torch.float32)
    args.make("velo_kalman_endvelo_states", STATE_FLOATS * nt, torch.float32)
    p = velo.KalmanParams()
    p.host_number_of_reconstructed_velo_tracks, p.host_number_of_events = nt, n
    _list_fields(p, args)
    p.dev_number_of_events = args.ptr("number_of_events")
    p.dev_offsets_all_velo_tracks = args.ptr("offsets_all_velo_tracks")
    if from_views:
        p.dev_velo_tracks_view = args.ptr("velo_tracks_view")
    else:
        p.dev_offsets_velo_track_hit_number = args.ptr("offsets_velo_track_hit_number")
        p.dev_velo_track_hits = args.ptr("velo_track_hits")
    p.dev_is_backward = args.ptr("is_backward")
    p.dev_velo_kalman_beamline_states = args.ptr("velo_kalman_beamline_states")
    p.dev_velo_kalman_endvelo_states = args.ptr("velo_kalman_endvelo_states")
    for name in ["velo_kalman_beamline_states_view", "velo_kalman_endvelo_states_view"]:
        args.make(name, n * velo.KALMAN_STATES_VIEW_DTYPE.itemsize, torch.uint8)
        setattr(p, "dev_" + name, args.ptr(name))
    p.constants, p.covariance_fp64_fallback = constants.handle, covariance_fp64_fallback
    velo.check(velo.lib().velo_kalman_filter(ctypes.byref(p), None), "velo_kalman_filter")


def velo_sp_sequence(constants, events, number_of_events, event_list=None, event_start=0, poison=False,
                     **props) -> Arguments:
    """The nine algorithms and five host prefix sums of the VELO-SP sequence, in the reference's
    order (velo_reconstruction.py:72-321), over `number_of_events` events starting at raw event
    `event_start`, visiting `event_list` (default: all of them). Returns the argument store."""
    torch = velo._torch()
    args = Arguments(events.raw.device, poison)
    n = int(number_of_events)
    el = np.arange(n, dtype=np.uint32) if event_list is None else np.asarray(event_list, dtype=np.uint32)
    if el.size and int(el.max()) >= n:
        raise velo.VeloError("event list entries must be below the number of events")
    args["event_list"] = torch.from_numpy(el.view(np.int32).copy()).to(args.device)
    args["number_of_events"] = torch.tensor([n], dtype=torch.int32, device=args.device)
    args.host.update(number_of_events=n, event_list_size=int(el.size), event_start=int(event_start))

    velo_calculate_number_of_candidates_t(args, events, props.get("block_dim_x", 256))
    args.host["number_of_cluster_candidates"] = host_prefix_sum(args, "number_of_candidates", "candidates_offsets", n)
    velo_estimate_input_size_t(args, events)
    args.host["total_number_of_velo_clusters"] = host_prefix_sum(
        args, "estimated_input_size", "offsets_estimated_input_size", N_MODULE_PAIRS * n)
    velo_masked_clustering_t(args, events, constants)
    velo_sort_by_phi_t(args)
    velo_search_by_triplet_t(args, constants, props.get("phi_tolerance", 0.045), props.get("max_scatter", 0.08),
                             props.get("max_skipped_modules", 1), props.get("search_block", 0))
    args.host["number_of_velo_tracks_at_least_four_hits"] = host_prefix_sum(
        args, "number_of_velo_tracks", "offsets_velo_tracks", n)
    velo_three_hit_tracks_filter_t(args, props.get("max_chi2", 20.0))
    args.host["number_of_three_hit_tracks_filtered"] = host_prefix_sum(
        args, "number_of_three_hit_tracks_output", "offsets_three_hit_tracks_filtered", n)
    nt = velo_copy_track_hit_number_t(args)
    args.host["number_of_reconstructed_velo_tracks"] = nt
    args.host["accumulated_number_of_hits_in_velo_tracks"] = host_prefix_sum(
        args, "velo_track_hit_number", "offsets_velo_track_hit_number", nt)
    velo_consolidate_tracks_t(args)
    velo_kalman_filter_t(args, constants, props.get("kalman_fp64_fallback", 0), props.get("kalman_from_views", True))
    return args
