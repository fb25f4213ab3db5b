"""The MEP layout (main/include/MEPTools.h:59-200): the generator's converter and the oracle's MEP
reader (runs without a GPU).

The same events in Allen's layout and in the MEP layout must decode to the same outputs: every
buffer of the oracle's sequence is compared. The converter is checked against the accessors'
definitions on its own: each fragment is the Allen bank's bytes after the source ID, and the sizes
and types arrays are indexed by bank, then event.
"""
import os

import numpy as np
import pytest

import oracle
from allen_amd import synth
from tests import textbook

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

CASES = {
    "v4_pileup": dict(n_events=12, pileup=1, nu=7.6, seed=91),
    "v2_single_pv": dict(n_events=8, pileup=0, mean_tracks=120.0, bank_version=2, seed=92),
    "retina_v3": dict(n_events=8, pileup=1, nu=7.6, retina=1, bank_version=3, seed=93),
    "retina_v4": dict(n_events=8, pileup=1, nu=7.6, retina=1, bank_version=4, seed=94),
}


def _events(name):
    if name == "mdf_v2":
        return textbook.load_mdf_fixture(os.path.join(GOLDEN, "mdf_minbias_vp.npz"))
    return synth.generate(**CASES[name])


def test_converter_against_accessors():
    ev = synth.generate(n_events=4, pileup=1, nu=7.6, seed=95, bank_version=2)
    m = synth.to_mep(ev)
    nb = int(m.offsets[0])
    assert m.mep and int(m.offsets[1]) == ev.n_events
    s16, t8 = ev.sizes.view("<u2"), ev.types.view(np.uint8)
    ms16, mt8 = m.sizes.view("<u2"), m.types.view(np.uint8)
    for e in range(ev.n_events):
        base = int(ev.offsets[e])
        assert int(ev.raw[base:base + 4].view("<u4")[0]) == nb
        boff = ev.raw[base + 4:base + 8 + 4 * nb].view("<u4")
        payload = base + 8 + 4 * nb
        for b in range(nb):
            bank = ev.raw[payload + int(boff[b]):payload + int(boff[b + 1])]
            assert int(bank[:4].view("<u4")[0]) == int(m.offsets[2 + b])              # MEP::source_id
            frag = int(m.offsets[2 + nb * (1 + e) + b])                                # MEP::offset_index
            assert np.array_equal(m.raw[frag:frag + bank.size - 4], bank[4:])
            assert ms16[int(m.sizes[b]) + e] == s16[int(ev.sizes[e]) + b]              # MEP::bank_size
            assert mt8[int(m.types[b]) + e] == t8[int(ev.types[e]) + b]                # MEP::bank_type


@pytest.mark.parametrize("name", list(CASES) + ["mdf_v2"])
def test_oracle_reads_mep_like_allen(name, geometry, beamline):
    ev = _events(name)
    a = oracle.run(ev, geometry, beamline, compute_f64=0)
    b = oracle.run(synth.to_mep(ev), geometry, beamline, compute_f64=0)
    names = ["module_cluster_num", "offsets_estimated_input_size", "sorted_velo_cluster_container",
             "offsets_all_velo_tracks", "velo_track_hits", "velo_kalman_beamline_states"]
    if not getattr(ev, "retina", False):
        names += ["number_of_candidates", "velo_bank_index", "estimated_input_size", "cluster_candidates"]
    for k in names:
        assert np.array_equal(a[k], b[k]), k
    assert int(a["offsets_all_velo_tracks"][-1]) > 0


def test_event_start_and_list_in_mep(geometry, beamline):
    ev = synth.generate(n_events=10, pileup=1, nu=7.6, seed=96)
    kw = dict(n_events=6, event_start=3, event_list=[5, 1, 4], compute_f64=0)
    a = oracle.run(ev, geometry, beamline, **kw)
    b = oracle.run(synth.to_mep(ev), geometry, beamline, **kw)
    for k in ["module_cluster_num", "offsets_all_velo_tracks", "velo_track_hits"]:
        assert np.array_equal(a[k], b[k]), k
