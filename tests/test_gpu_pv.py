"""The beamline PV finder on the GPU against the CPU oracle (runs on an MI355X).

Exact: the PV tracks, the z histograms, the seeds. Within tolerance: the track denominators,
vertices and their covariances, which go through expf (a few ulp apart between the host's libm and
the GPU's), stated below."""
import ctypes

import numpy as np
import pytest

import oracle
from allen_amd import synth, velo

pytestmark = pytest.mark.gpu

DATASETS = {
    "pileup": dict(n_events=80, pileup=1, nu=7.6, seed=91),
    "single_pv": dict(n_events=40, pileup=0, mean_tracks=120.0, seed=92),
    "stress": dict(n_events=16, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01, seed=93),
    "retina": dict(n_events=40, pileup=1, nu=7.6, seed=94, retina=1),
    "empty": dict(n_events=6, pileup=0, mean_tracks=0.0, noise_fraction=0.0, seed=95),
}
V = velo.PV_VERTEX_DTYPE


@pytest.fixture(scope="module")
def constants(geometry, beamline):
    c = velo.VeloConstants(geometry, beamline)
    yield c
    c.close()


def _compare(gpu, ref, n, nt):
    np.testing.assert_array_equal(gpu["pvtracks"].view(np.uint32), ref["pvtracks"][:7 * nt].view(np.uint32))
    if "zhisto" in gpu:
        np.testing.assert_array_equal(gpu["zhisto"].view(np.uint32), ref["zhisto"].view(np.uint32))
    np.testing.assert_array_equal(gpu["number_of_zpeaks"], ref["number_of_zpeaks"])
    nz = ref["number_of_zpeaks"]
    zp_g, zp_r = gpu["zpeaks"].reshape(n, 32), ref["zpeaks"].reshape(n, 32)
    for e in range(n):
        np.testing.assert_array_equal(zp_g[e, :nz[e]].view(np.uint32), zp_r[e, :nz[e]].view(np.uint32))
    np.testing.assert_allclose(gpu["pvtracks_denom"], ref["pvtracks_denom"][:nt], rtol=1e-5, atol=1e-30)
    for kind in ("multi_fit", "multi_final"):
        cnt = f"number_of_{kind}_vertices"
        np.testing.assert_array_equal(gpu[cnt], ref[cnt])
        g = gpu[f"{kind}_vertices"].reshape(n, 32)
        r = ref[f"{kind}_vertices"].reshape(n, 32)
        for e in range(n):
            a, b = g[e, :ref[cnt][e]], r[e, :ref[cnt][e]]
            for f in ("x", "y", "z"):
                np.testing.assert_allclose(a[f], b[f], rtol=0, atol=2e-5, err_msg=f"{kind} {f} event {e}")
            for f in ("cov00", "cov11", "cov22", "cov10", "cov20", "cov21", "chi2", "nTracks"):
                np.testing.assert_allclose(a[f], b[f], rtol=1e-4, atol=1e-9, err_msg=f"{kind} {f} event {e}")


@pytest.mark.parametrize("name", list(DATASETS))
def test_pv_finder_matches_oracle(name, geometry, beamline, constants):
    ev = synth.generate(**DATASETS[name])
    print(f"[pv] dataset={name} events={ev.n_events}", flush=True)
    ref = oracle.run(ev, geometry, beamline, pv_finder=1)
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, 4096))
    seq.set_tuning("pv_finder", 2)
    seq.run(velo.DeviceEvents(ev))
    assert seq.status() & (velo.STATUS_OVERFLOW_MASK | (1 << 16)) == 0
    n = ev.n_events
    nt = int(ref["offsets_all_velo_tracks"][-1])
    gpu = {"pvtracks": seq.fetch("pvtracks", np.float32, 7 * nt), "zhisto": seq.fetch("zhisto", np.float32, 3392 * n),
           "zpeaks": seq.fetch("zpeaks", np.float32, 32 * n),
           "number_of_zpeaks": seq.fetch("number_of_zpeaks", np.uint32, n),
           "pvtracks_denom": seq.fetch("pvtracks_denom", np.float32, nt)}
    for kind in ("multi_fit", "multi_final"):
        gpu[f"{kind}_vertices"] = seq.fetch(f"{kind}_vertices", V, 32 * n)
        gpu[f"number_of_{kind}_vertices"] = seq.fetch(f"number_of_{kind}_vertices", np.uint32, n)
    seq.close()
    _compare(gpu, ref, n, nt)
    if name != "empty":
        assert ref["number_of_multi_final_vertices"].sum() > 0


CROWDED = {
    "nu60": dict(n_events=8, pileup=1, nu=60.0, pv_sigma_z=150.0, seed=96),
    "nu120": dict(n_events=6, pileup=1, nu=120.0, pv_sigma_z=200.0, seed=97),
}


@pytest.mark.parametrize("name", list(CROWDED))
def test_pv_finder_seed_overflow_matches_oracle(name, geometry, beamline, constants):
    """Events with more seeds than the 32 kept (and, at nu 120, more tracks than one histogram
    chunk and many multi-fitter chunks): the first 32 seeds and their vertices as the oracle's,
    and the same overflow bit."""
    ev = synth.generate(**CROWDED[name])
    print(f"[pv] crowded dataset={name} events={ev.n_events}", flush=True)
    ref = oracle.run(ev, geometry, beamline, pv_finder=1)
    assert int(ref["status"][0]) == 1 << 16, "the oracle should flag the seed overflow and nothing else"
    seq = velo.VeloSequence(constants, ev.n_events, max(ev.raw.size, 4096))
    seq.set_tuning("pv_finder", 2)
    seq.run(velo.DeviceEvents(ev))
    assert seq.status() & ~velo.STATUS_LDS_FALLBACK == 1 << 16  # crowded events may take the global-memory path
    n = ev.n_events
    nt = int(ref["offsets_all_velo_tracks"][-1])
    gpu = {"pvtracks": seq.fetch("pvtracks", np.float32, 7 * nt), "zhisto": seq.fetch("zhisto", np.float32, 3392 * n),
           "zpeaks": seq.fetch("zpeaks", np.float32, 32 * n),
           "number_of_zpeaks": seq.fetch("number_of_zpeaks", np.uint32, n),
           "pvtracks_denom": seq.fetch("pvtracks_denom", np.float32, nt)}
    for kind in ("multi_fit", "multi_final"):
        gpu[f"{kind}_vertices"] = seq.fetch(f"{kind}_vertices", V, 32 * n)
        gpu[f"number_of_{kind}_vertices"] = seq.fetch(f"number_of_{kind}_vertices", np.uint32, n)
    seq.close()
    _compare(gpu, ref, n, nt)
    assert (ref["number_of_zpeaks"] == 32).sum() >= n // 2


def test_pv_drop_in_entry_point(geometry, beamline, constants):
    """velo_pv_beamline called through the C ABI on the sequence's beamline states, with an
    event list in reverse order, gives the oracle's vertices."""
    import torch
    ev = synth.generate(**DATASETS["pileup"])
    n = ev.n_events
    ref = oracle.run(ev, geometry, beamline, pv_finder=1)
    seq = velo.VeloSequence(constants, n, ev.raw.size)
    seq.run(velo.DeviceEvents(ev))
    nt = int(ref["offsets_all_velo_tracks"][-1])
    z = lambda k, dt=torch.float32: torch.zeros(k, dtype=dt, device="cuda")
    bufs = {"pvtracks": z(7 * nt), "zpeaks": z(32 * n), "number_of_zpeaks": z(n, torch.int32),
            "pvtracks_denom": z(nt), "fit": z(12 * 32 * n), "nfit": z(n, torch.int32), "fin": z(12 * 32 * n),
            "nfin": z(n, torch.int32), "status": z(1, torch.int32)}
    event_list = torch.arange(n - 1, -1, -1, dtype=torch.int32, device="cuda")
    n_dev = torch.tensor([n], dtype=torch.int32, device="cuda")
    p = velo.PVParams()
    p.host_number_of_events, p.dev_event_list, p.event_list_size = n, event_list.data_ptr(), n
    p.dev_number_of_events = n_dev.data_ptr()
    p.dev_offsets_all_velo_tracks = seq.buffer_ptr("offsets_all_velo_tracks")[0]
    p.dev_velo_kalman_beamline_states = seq.buffer_ptr("velo_kalman_beamline_states")[0]
    p.dev_pvtracks, p.dev_zhisto, p.dev_zpeaks = bufs["pvtracks"].data_ptr(), 0, bufs["zpeaks"].data_ptr()
    p.dev_number_of_zpeaks, p.dev_pvtracks_denom = bufs["number_of_zpeaks"].data_ptr(), bufs["pvtracks_denom"].data_ptr()
    p.dev_multi_fit_vertices, p.dev_number_of_multi_fit_vertices = bufs["fit"].data_ptr(), bufs["nfit"].data_ptr()
    p.dev_multi_final_vertices = bufs["fin"].data_ptr()
    p.dev_number_of_multi_final_vertices = bufs["nfin"].data_ptr()
    p.dev_status, p.constants = bufs["status"].data_ptr(), constants.handle
    velo.check(velo.lib().velo_pv_beamline(ctypes.byref(p), None), "velo_pv_beamline")
    torch.cuda.synchronize()
    host = lambda t, dt: t.cpu().numpy().view(dt)
    gpu = {"pvtracks": host(bufs["pvtracks"], np.float32), "zpeaks": host(bufs["zpeaks"], np.float32),
           "number_of_zpeaks": host(bufs["number_of_zpeaks"], np.uint32),
           "pvtracks_denom": host(bufs["pvtracks_denom"], np.float32),
           "multi_fit_vertices": host(bufs["fit"], V), "number_of_multi_fit_vertices": host(bufs["nfit"], np.uint32),
           "multi_final_vertices": host(bufs["fin"], V),
           "number_of_multi_final_vertices": host(bufs["nfin"], np.uint32)}
    seq.close()
    assert int(host(bufs["status"], np.uint32)[0]) == 0
    _compare(gpu, ref, n, nt)
