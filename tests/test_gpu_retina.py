"""Retina-cluster decoding on the GPU against the CPU oracle (runs on an MI355X)."""
import ctypes

import numpy as np
import pytest

import oracle
from allen_amd import synth, velo
from tests import parity, textbook

pytestmark = pytest.mark.gpu

DATASETS = {
    "v4_pileup": dict(n_events=60, pileup=1, nu=7.6, seed=51, retina=1, bank_version=4),
    "v3_single_pv": dict(n_events=30, pileup=0, mean_tracks=120.0, seed=52, retina=1, bank_version=3),
    "v4_shuffled_backward": dict(n_events=30, pileup=0, mean_tracks=150.0, backward_fraction=0.3, shuffle_sp=1,
                                 seed=53, retina=1, bank_version=4),
    "v4_stress": dict(n_events=12, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01, seed=54,
                      retina=1, bank_version=4),
    "v4_empty": dict(n_events=6, pileup=0, mean_tracks=0.0, noise_fraction=0.0, seed=55, retina=1, bank_version=4),
}


@pytest.fixture(scope="module")
def constants(geometry, beamline):
    c = velo.VeloConstants(geometry, beamline)
    yield c
    c.close()


def _events(name):
    if name == "mdf_real_v3":
        import os
        ev = textbook.load_mdf_fixture(os.path.join(os.path.dirname(__file__), "golden", "mdf_minbias_retina.npz"))
        ev.retina = True
        return ev
    return synth.generate(**DATASETS[name])


@pytest.mark.parametrize("name", list(DATASETS) + ["mdf_real_v3"])
def test_retina_sequence_matches_oracle(name, geometry, beamline, constants):
    """The whole path from retina banks: decoding (sizes, offsets, keys, permutations, sorted
    container) and everything after it equal the oracle's bit for bit."""
    ev = _events(name)
    print(f"[retina] dataset={name} events={ev.n_events} version={ev.bank_version} raw_bytes={ev.raw.size}", flush=True)
    ref = oracle.run(ev, geometry, beamline)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, 4096))
    seq.run(velo.DeviceEvents(ev))
    assert seq.status() & velo.STATUS_OVERFLOW_MASK == 0
    gpu = parity.gpu_buffers(seq, ref)
    seq.close()
    errs = parity.compare(gpu, ref, True)
    assert not errs, "\n".join(errs[:20])
    if name != "v4_empty":
        assert ref["offsets_all_velo_tracks"][-1] > 0


def test_retina_drop_in_entry_points(geometry, beamline, constants):
    """calculate_number_of_retinaclusters_each_sensor_pair and decode_retinaclusters called
    one by one through the C ABI, with an event list, equal the oracle."""
    import torch
    ev = _events("v4_pileup")
    print(f"[retina-abi] events={ev.n_events}", flush=True)
    ref = oracle.run(ev, geometry, beamline)
    dev = velo.DeviceEvents(ev)
    n = ev.n_events
    slots = velo.lib().velo_retina_sensor_slots(ev.bank_version)
    cuda = lambda k, dt=torch.int32: torch.zeros(k, dtype=dt, device="cuda")
    event_list = torch.arange(n, dtype=torch.int32, device="cuda")
    sizes, bank_index = cuda(n * slots), cuda(slots)
    p = velo.RetinaSizesParams()
    p.host_number_of_events, p.host_raw_bank_version = n, ev.bank_version
    p.dev_event_list, p.event_list_size = event_list.data_ptr(), n
    p.dev_velo_retina_raw_input, p.dev_velo_retina_raw_input_offsets = dev.raw.data_ptr(), dev.offsets.data_ptr()
    p.dev_velo_retina_raw_input_sizes, p.dev_velo_retina_raw_input_types = dev.sizes.data_ptr(), dev.types.data_ptr()
    p.dev_retina_bank_index, p.dev_each_sensor_pair_size = bank_index.data_ptr(), sizes.data_ptr()
    velo.check(velo.lib().velo_calculate_number_of_retinaclusters_each_sensor_pair(ctypes.byref(p), None), "sizes")
    offsets = cuda(n * slots + 1)
    total = velo.prefix_sum(sizes, n * slots, offsets, want_total=True)
    assert total == int(ref["offsets_each_sensor_pair_size"][-1])
    n_events_dev = torch.tensor([n], dtype=torch.int32, device="cuda")
    mcn, mp_off = cuda(26 * n), cuda(26 * n + 1)
    container, perm, keys = cuda(18 * total, torch.uint8), cuda(total), cuda(total, torch.int64)
    d = velo.RetinaDecodeParams()
    d.host_total_number_of_velo_clusters, d.host_raw_bank_version, d.host_number_of_events = total, ev.bank_version, n
    d.dev_velo_retina_raw_input, d.dev_velo_retina_raw_input_offsets = dev.raw.data_ptr(), dev.offsets.data_ptr()
    d.dev_velo_retina_raw_input_sizes, d.dev_velo_retina_raw_input_types = dev.sizes.data_ptr(), dev.types.data_ptr()
    d.dev_offsets_each_sensor_pair_size, d.dev_retina_bank_index = offsets.data_ptr(), bank_index.data_ptr()
    d.dev_event_list, d.event_list_size, d.dev_number_of_events = event_list.data_ptr(), n, n_events_dev.data_ptr()
    d.dev_module_cluster_num, d.dev_offsets_module_pair_cluster = mcn.data_ptr(), mp_off.data_ptr()
    d.dev_velo_cluster_container, d.dev_hit_permutations = container.data_ptr(), perm.data_ptr()
    d.dev_hit_sorting_key, d.constants = keys.data_ptr(), constants.handle
    velo.check(velo.lib().velo_decode_retinaclusters(ctypes.byref(d), None), "decode")
    torch.cuda.synchronize()
    host = lambda t, dt: t.cpu().numpy().view(dt)
    np.testing.assert_array_equal(host(sizes, np.uint32), ref["each_sensor_pair_size"])
    np.testing.assert_array_equal(host(bank_index, np.uint32), ref["retina_bank_index"])
    np.testing.assert_array_equal(host(mcn, np.uint32), ref["module_cluster_num"])
    np.testing.assert_array_equal(host(mp_off, np.uint32), ref["offsets_estimated_input_size"])
    np.testing.assert_array_equal(host(keys, np.int64), ref["hit_sorting_key"])
    np.testing.assert_array_equal(host(perm, np.uint32), ref["hit_permutations"])
    slots_f = parity.filled_slots(ref["offsets_estimated_input_size"], ref["module_cluster_num"])
    g = parity.cluster_views(host(container, np.uint8), total)
    r = parity.cluster_views(ref["sorted_velo_cluster_container"], total)
    for a, b in zip(g, r):
        np.testing.assert_array_equal(a[slots_f].view(np.uint8), b[slots_f].view(np.uint8))
