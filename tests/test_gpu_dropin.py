"""The nine VELO-SP algorithms called one by one through the C ABI, as an Allen sequence binds them
(velo_reconstruction.py:72-321), against the CPU oracle (runs on an MI355X).

The event list is a reversed subset (Allen's event filters hand the kernels such lists) and the
raw input holds more events than the run covers, with event_start != 0, so raw event e of the run
is raw block e + event_start. Output buffers start as all-ones bytes: every value the reference
writes or clears for an event outside the list has to be written by the library too.
"""
import numpy as np
import pytest

import oracle
from allen_amd import algorithms, synth, velo
from tests import parity

pytestmark = pytest.mark.gpu

CASES = {
    # name: (generator settings, run events, event_start, event list)
    "v4_odd_reversed": (dict(n_events=40, pileup=1, nu=7.6, seed=61), 30, 6, list(range(29, 0, -2))),
    "v4_full_shifted": (dict(n_events=24, pileup=0, mean_tracks=120.0, seed=62), 20, 4, list(range(20))),
    "v2_subset": (dict(n_events=16, pileup=0, mean_tracks=100.0, bank_version=2, seed=63), 12, 3, [7, 2, 11, 5]),
    "empty_list": (dict(n_events=8, pileup=1, nu=7.6, seed=64), 6, 1, []),
    # MEP layout (mep_layout = 1): the same kind of events as fragments grouped by bank.
    "mep_v4_subset": (dict(n_events=30, pileup=1, nu=7.6, seed=67), 24, 5, [20, 3, 11, 7, 0]),
    "mep_v2_subset": (dict(n_events=16, pileup=0, mean_tracks=100.0, bank_version=2, seed=68), 12, 2, [9, 4, 6]),
}


@pytest.fixture(scope="module")
def constants(geometry, beamline):
    c = velo.VeloConstants(geometry, beamline)
    yield c
    c.close()


def _host(args, name, dtype, count):
    a = args[name].cpu().numpy().view(np.uint8).view(dtype)
    return a[:int(count)].copy()


def gpu_store(args, ref) -> dict:
    """The drop-in outputs under the oracle's names, cut to the oracle's sizes."""
    n = ref["number_of_velo_tracks"].size
    total = int(ref["offsets_estimated_input_size"][-1])
    nt = int(ref["offsets_all_velo_tracks"][-1])
    out = {}
    for name in ["number_of_candidates", "module_candidate_num", "number_of_velo_tracks",
                 "number_of_three_hit_tracks_output"]:
        out[name] = _host(args, name, np.uint32, n)
    for name in ["candidates_offsets", "offsets_velo_tracks", "offsets_three_hit_tracks_filtered",
                 "offsets_all_velo_tracks"]:
        out[name] = _host(args, name, np.uint32, n + 1)
    out["velo_bank_index"] = _host(args, "velo_bank_index", np.uint32, 208)
    out["estimated_input_size"] = _host(args, "estimated_input_size", np.uint32, 26 * n)
    out["module_cluster_num"] = _host(args, "module_cluster_num", np.uint32, 26 * n)
    out["offsets_estimated_input_size"] = _host(args, "offsets_estimated_input_size", np.uint32, 26 * n + 1)
    out["cluster_candidates"] = _host(args, "cluster_candidates", np.uint32, ref["cluster_candidates"].size)
    out["velo_cluster_container"] = _host(args, "velo_cluster_container", np.uint8, 18 * total)
    out["sorted_velo_cluster_container"] = _host(args, "sorted_velo_cluster_container", np.uint8, 18 * total)
    out["hit_permutation"] = _host(args, "hit_permutation", np.uint32, total)
    out["tracks"] = _host(args, "tracks", parity.TRACK, ref["tracks"].size)
    out["tracklets"] = _host(args, "tracklets", parity.TRACKLET, ref["tracklets"].size)
    out["tracks_to_follow"] = _host(args, "tracks_to_follow", np.uint32, ref["tracks_to_follow"].size)
    out["three_hit_tracks"] = _host(args, "three_hit_tracks", parity.TRACKLET, ref["three_hit_tracks"].size)
    out["three_hit_tracks_output"] = _host(args, "three_hit_tracks_output", parity.TRACKLET,
                                           ref["three_hit_tracks_output"].size)
    out["hit_used"] = _host(args, "hit_used", np.bool_, total)
    out["rel_indices"] = _host(args, "rel_indices", np.uint16, total)
    out["atomics_velo"] = _host(args, "atomics_velo", np.uint32, 3 * n)
    out["velo_track_hit_number"] = _host(args, "velo_track_hit_number", np.uint32, nt)
    out["offsets_velo_track_hit_number"] = _host(args, "offsets_velo_track_hit_number", np.uint32, nt + 1)
    out["velo_track_hits"] = _host(args, "velo_track_hits", np.uint8, ref["velo_track_hits"].size)
    out["is_backward"] = _host(args, "is_backward", np.bool_, nt)
    out["accepted_velo_tracks"] = _host(args, "accepted_velo_tracks", np.uint8, nt)
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        out[name] = _host(args, name, np.float32, 14 * nt)
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_sp_drop_ins_with_event_list(case, geometry, beamline, constants):
    settings, n_run, event_start, event_list = CASES[case]
    ev = synth.generate(**settings)
    if case.startswith("mep_"):
        ev = synth.to_mep(ev)
    print(f"[drop-in] case={case} raw_events={ev.n_events} run={n_run} start={event_start} "
          f"list={len(event_list)}", flush=True)
    assert n_run + event_start <= ev.n_events
    ref = oracle.run(ev, geometry, beamline, n_events=n_run, event_start=event_start, event_list=event_list)
    dev = velo.DeviceEvents(ev)
    args = algorithms.velo_sp_sequence(constants, dev, n_run, event_list=event_list, event_start=event_start,
                                       poison=True)
    assert int(args["search_status"].cpu()[0]) & velo.STATUS_OVERFLOW_MASK == 0
    gpu = gpu_store(args, ref)
    errs = parity.compare(gpu, ref, fused=False)
    assert not errs, "\n".join(errs[:20])
    # Events left out of the list own nothing downstream.
    unlisted = np.setdiff1d(np.arange(n_run), np.asarray(event_list, dtype=np.int64))
    counts = np.diff(gpu["offsets_all_velo_tracks"].astype(np.int64))
    assert np.all(gpu["number_of_velo_tracks"][unlisted] == 0)
    assert np.all(gpu["atomics_velo"].reshape(-1, 3)[unlisted] == 0)
    assert np.all(counts[unlisted] == 0)
    assert np.all(gpu["estimated_input_size"].reshape(-1, 26)[unlisted] == 0)
    if len(event_list):
        assert ref["offsets_all_velo_tracks"][-1] > 0
    # The host totals the sequence read back equal the oracle's.
    assert args.host["number_of_reconstructed_velo_tracks"] == int(ref["offsets_all_velo_tracks"][-1])
    assert args.host["total_number_of_velo_clusters"] == int(ref["offsets_estimated_input_size"][-1])
    check_views(args, ref, event_list)


def _view(args, name, dtype, count):
    return args[name].cpu().numpy()[:dtype.itemsize * count].view(dtype)


def check_views(args, ref, event_list):
    """The view objects against the layouts Allen's constructors give them, built here from the
    oracle's offsets and the buffers' device addresses (VeloEventModel.cuh:91-110,
    VeloConsolidateTracks.cu:11-46, VeloConsolidated.cuh:60-294, States.cuh:191-220)."""
    n = ref["number_of_velo_tracks"].size
    oat = ref["offsets_all_velo_tracks"].astype(np.int64)
    ohn = ref["offsets_velo_track_hit_number"].astype(np.int64)
    nt = int(oat[-1])
    listed = np.asarray(event_list, dtype=np.int64)
    cl = _view(args, "velo_clusters", velo.CLUSTERS_VIEW_DTYPE, n)[listed]
    assert np.all(cl["base_pointer"].astype(np.int64) == args.ptr("velo_cluster_container"))
    assert np.all(cl["total_number_of_hits"] == int(ref["offsets_estimated_input_size"][-1]))
    assert np.all(cl["offset"] == 0)
    hits = _view(args, "velo_hits_view", velo.HITS_VIEW_DTYPE, n)
    assert np.all(hits["base_pointer"].astype(np.int64) == args.ptr("velo_track_hits"))
    np.testing.assert_array_equal(hits["offset"], ohn[oat[:-1]])
    np.testing.assert_array_equal(hits["size"], ohn[oat[1:]] - ohn[oat[:-1]])
    assert np.all(hits["total_number_of_hits"] == ohn[nt]) and np.all(hits["pad_"] == 0)
    tr = _view(args, "velo_track_view", velo.TRACK_VIEW_DTYPE, nt)
    ev = np.repeat(np.arange(n), np.diff(oat))
    t = np.arange(nt)
    np.testing.assert_array_equal(tr["hits"].astype(np.int64), args.ptr("velo_hits_view") + velo.HITS_VIEW_DTYPE.itemsize * ev)
    np.testing.assert_array_equal(tr["track_index"], t - oat[ev])
    np.testing.assert_array_equal(tr["track_container_offset"], oat[ev])
    np.testing.assert_array_equal(tr["offset"], ohn[t] - ohn[oat[ev]])
    np.testing.assert_array_equal(tr["number_of_hits"], np.diff(ohn[:nt + 1]))
    ts = _view(args, "velo_tracks_view", velo.TRACKS_VIEW_DTYPE, n)
    np.testing.assert_array_equal(ts["track"].astype(np.int64), args.ptr("velo_track_view") + velo.TRACK_VIEW_DTYPE.itemsize * oat[:-1])
    np.testing.assert_array_equal(ts["size"], np.diff(oat))
    np.testing.assert_array_equal(ts["offset"], oat[:-1])
    m = _view(args, "velo_multi_event_tracks_view", velo.MULTI_EVENT_TRACKS_VIEW_DTYPE, 1)[0]
    assert m["type_id"] == velo.TYPE_ID_VELO_TRACKS and m["container"].astype(np.int64) == args.ptr("velo_tracks_view")
    assert m["number_of_events"] == n and m["pad_"] == 0 and m["pad2_"] == 0
    assert args["imec_velo_tracks"].cpu().numpy()[:8].view("<u8")[0] == args.ptr("velo_multi_event_tracks_view")
    for name in ["velo_kalman_beamline_states", "velo_kalman_endvelo_states"]:
        sv = _view(args, name + "_view", velo.KALMAN_STATES_VIEW_DTYPE, n)[listed]
        assert np.all(sv["base_pointer"].astype(np.int64) == args.ptr(name))
        np.testing.assert_array_equal(sv["offset"], oat[listed])
        np.testing.assert_array_equal(sv["size"], oat[listed + 1] - oat[listed])
        assert np.all(sv["total_number_of_tracks"] == nt) and np.all(sv["pad_"] == 0)


def test_kalman_through_offsets_equals_views(geometry, beamline, constants):
    """The Kalman drop-in reading tracks through dev_velo_tracks_view (the reference's input) and
    through the hit-number offsets gives the same bytes."""
    ev = synth.generate(n_events=20, pileup=1, nu=7.6, seed=66)
    dev = velo.DeviceEvents(ev)
    out = {}
    for views in (True, False):
        args = algorithms.velo_sp_sequence(constants, dev, ev.n_events, event_list=list(range(19, -1, -3)),
                                           poison=True, kalman_from_views=views)
        nt = args.host["number_of_reconstructed_velo_tracks"]
        out[views] = [_host(args, k, np.uint8, n) for k, n in
                      [("velo_kalman_beamline_states", 56 * nt), ("velo_kalman_endvelo_states", 56 * nt),
                       ("is_backward", nt)]]
    for a, b in zip(out[True], out[False]):
        np.testing.assert_array_equal(a, b)


def test_retina_drop_ins_with_event_list(geometry, beamline, constants):
    """decode_retinaclusters with a subset list and event_start: module-pair offsets cover every
    event of the run (DecodeRetinaClusters.cu:11-36), listed or not."""
    import ctypes
    import torch
    ev = synth.generate(n_events=30, pileup=1, nu=7.6, seed=65, retina=1, bank_version=4)
    n_run, start, event_list = 24, 5, [22, 17, 3, 9, 0]
    print(f"[drop-in] retina run={n_run} start={start} list={event_list}", flush=True)
    ref = oracle.run(ev, geometry, beamline, n_events=n_run, event_start=start, event_list=event_list)
    dev = velo.DeviceEvents(ev)
    slots = velo.lib().velo_retina_sensor_slots(4)
    ones = lambda k, dt=torch.int32: torch.full((max(k, 1),), -1, dtype=dt, device="cuda")
    el = torch.tensor(event_list, dtype=torch.int32, device="cuda")
    sizes, bank_index = ones(n_run * slots), torch.zeros(slots, dtype=torch.int32, device="cuda")
    p = velo.RetinaSizesParams()
    p.host_number_of_events, p.host_raw_bank_version = n_run, 4
    p.dev_event_list, p.event_list_size = el.data_ptr(), len(event_list)
    p.dev_velo_retina_raw_input, p.dev_velo_retina_raw_input_offsets = dev.raw.data_ptr(), dev.offsets.data_ptr()
    p.dev_velo_retina_raw_input_sizes, p.dev_velo_retina_raw_input_types = dev.sizes.data_ptr(), dev.types.data_ptr()
    p.dev_retina_bank_index, p.dev_each_sensor_pair_size = bank_index.data_ptr(), sizes.data_ptr()
    p.event_start = start
    velo.check(velo.lib().velo_calculate_number_of_retinaclusters_each_sensor_pair(ctypes.byref(p), None), "sizes")
    offsets = ones(n_run * slots + 1)
    total = velo.prefix_sum(sizes, n_run * slots, offsets, want_total=True)
    assert total == int(ref["offsets_each_sensor_pair_size"][-1])
    n_events_dev = torch.tensor([n_run], dtype=torch.int32, device="cuda")
    mcn, mp_off = ones(26 * n_run), ones(26 * n_run + 1)
    container, perm, keys = ones(18 * total, torch.uint8), ones(total), ones(total, torch.int64)
    d = velo.RetinaDecodeParams()
    d.host_total_number_of_velo_clusters, d.host_raw_bank_version, d.host_number_of_events = total, 4, n_run
    d.dev_velo_retina_raw_input, d.dev_velo_retina_raw_input_offsets = dev.raw.data_ptr(), dev.offsets.data_ptr()
    d.dev_velo_retina_raw_input_sizes, d.dev_velo_retina_raw_input_types = dev.sizes.data_ptr(), dev.types.data_ptr()
    d.dev_offsets_each_sensor_pair_size, d.dev_retina_bank_index = offsets.data_ptr(), bank_index.data_ptr()
    d.dev_event_list, d.event_list_size, d.dev_number_of_events = el.data_ptr(), len(event_list), n_events_dev.data_ptr()
    d.dev_module_cluster_num, d.dev_offsets_module_pair_cluster = mcn.data_ptr(), mp_off.data_ptr()
    d.dev_velo_cluster_container, d.dev_hit_permutations = container.data_ptr(), perm.data_ptr()
    d.dev_hit_sorting_key, d.constants, d.event_start = keys.data_ptr(), constants.handle, start
    views = ones(n_run * velo.CLUSTERS_VIEW_DTYPE.itemsize, torch.uint8)
    d.dev_velo_clusters = views.data_ptr()
    velo.check(velo.lib().velo_decode_retinaclusters(ctypes.byref(d), None), "decode")
    torch.cuda.synchronize()
    host = lambda t, dt, k: t.cpu().numpy().view(dt)[:k]
    np.testing.assert_array_equal(host(sizes, np.uint32, n_run * slots), ref["each_sensor_pair_size"])
    np.testing.assert_array_equal(host(bank_index, np.uint32, slots), ref["retina_bank_index"])
    np.testing.assert_array_equal(host(mcn, np.uint32, 26 * n_run), ref["module_cluster_num"])
    np.testing.assert_array_equal(host(mp_off, np.uint32, 26 * n_run + 1), ref["offsets_estimated_input_size"])
    np.testing.assert_array_equal(host(keys, np.int64, total), ref["hit_sorting_key"])
    np.testing.assert_array_equal(host(perm, np.uint32, total), ref["hit_permutations"])
    slots_f = parity.filled_slots(ref["offsets_estimated_input_size"], ref["module_cluster_num"])
    g = parity.cluster_views(host(container, np.uint8, 18 * total), total)
    r = parity.cluster_views(ref["sorted_velo_cluster_container"], total)
    for a, b in zip(g, r):
        np.testing.assert_array_equal(a[slots_f].view(np.uint8), b[slots_f].view(np.uint8))
    # Velo::Clusters views (DecodeRetinaClusters.cu:331-346,442): zeroed, then the listed events'.
    v = views.cpu().numpy().view(velo.CLUSTERS_VIEW_DTYPE)
    listed = np.zeros(n_run, bool)
    listed[event_list] = True
    assert np.all(v["base_pointer"][listed].astype(np.int64) == container.data_ptr())
    assert np.all(v["total_number_of_hits"][listed] == total) and np.all(v["offset"] == 0)
    assert np.all(v["base_pointer"][~listed] == 0) and np.all(v["total_number_of_hits"][~listed] == 0)
