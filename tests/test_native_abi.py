"""The C-ABI library loads and exports every function include/velo_amd.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

from allen_amd import velo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "velo_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(velo_[a-z0-9_]+)\s*\(", text))
    return names


def test_header_declares_the_entry_points():
    names = declared_functions()
    for required in ["velo_calculate_number_of_candidates", "velo_estimate_input_size", "velo_masked_clustering",
                     "velo_sort_by_phi", "velo_search_by_triplet", "velo_three_hit_tracks_filter",
                     "velo_copy_track_hit_number", "velo_consolidate_tracks", "velo_kalman_filter",
                     "velo_prefix_sum", "velo_cache_sp_patterns", "velo_sequence_run",
                     "velo_calculate_number_of_retinaclusters_each_sensor_pair", "velo_decode_retinaclusters"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(velo.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(velo.SIGNATURES) == declared_functions()
    velo.lib()  # binds every signature


def test_params_structs_match_header_layout():
    """Field order of the ctypes mirrors equals the C structs' (parsed from the header)."""
    text = open(os.path.join(ROOT, "include", "velo_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    structs = dict((m.group(2), m.group(1)) for m in re.finditer(r"typedef struct \{(.*?)\}\s*(\w+);", text, re.S))
    mirrors = {"velo_calculate_number_of_candidates_params": velo.CandidatesParams,
               "velo_estimate_input_size_params": velo.EstimateParams,
               "velo_masked_clustering_params": velo.MaskedClusteringParams,
               "velo_sort_by_phi_params": velo.SortParams,
               "velo_search_by_triplet_params": velo.SearchParams,
               "velo_three_hit_tracks_filter_params": velo.ThreeHitParams,
               "velo_copy_track_hit_number_params": velo.CopyHitNumberParams,
               "velo_consolidate_tracks_params": velo.ConsolidateParams,
               "velo_kalman_filter_params": velo.KalmanParams,
               "velo_prefix_sum_params": velo.PrefixSumParams,
               "velo_calculate_number_of_retinaclusters_params": velo.RetinaSizesParams,
               "velo_decode_retinaclusters_params": velo.RetinaDecodeParams,
               "velo_pv_beamline_params": velo.PVParams,
               "velo_sequence_input": velo.SequenceInput}
    for cname, py in mirrors.items():
        body = structs[cname]
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            for part in decl.split(","):
                fields.append(re.findall(r"(\w+)\s*$", part.strip())[0])
        assert [f[0] for f in py._fields_] == fields, cname


def test_bank_version_error_without_device():
    """A bad bank version is rejected before any device work (EstimateInputSize.cu:265-268)."""
    p = velo.CandidatesParams()
    p.host_raw_bank_version = 7
    p.dev_number_of_candidates = 1
    p.dev_velo_bank_index = 1
    assert velo.lib().velo_calculate_number_of_candidates(ctypes.byref(p), None) == 2


def test_cache_sp_patterns_host_only():
    pat, sizes, fx, fy = velo.cache_sp_patterns()
    assert pat.shape == (256,) and fx.shape == (512,)
    assert pat[0x01] == 0 and pat[0x81] & 8


def test_view_layouts_match_header():
    """The view structs of the header have the reference classes' sizes (Velo::Clusters 16 B, Hits
    24 B, Track 24 B, Tracks 16 B, MultiEventContainer<Tracks> 24 B, KalmanStates 24 B) and the
    Python dtypes name the header's fields in order."""
    text = open(os.path.join(ROOT, "include", "velo_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    structs = dict((m.group(2), m.group(1)) for m in re.finditer(r"typedef struct \{(.*?)\}\s*(\w+);", text, re.S))
    views = {"velo_clusters_view": (velo.CLUSTERS_VIEW_DTYPE, 16), "velo_hits_view": (velo.HITS_VIEW_DTYPE, 24),
             "velo_track_view": (velo.TRACK_VIEW_DTYPE, 24), "velo_tracks_view": (velo.TRACKS_VIEW_DTYPE, 16),
             "velo_multi_event_tracks_view": (velo.MULTI_EVENT_TRACKS_VIEW_DTYPE, 24),
             "velo_kalman_states_view": (velo.KALMAN_STATES_VIEW_DTYPE, 24)}
    for cname, (dt, size) in views.items():
        fields = [re.findall(r"(\w+)\s*$", d.strip())[0] for d in structs[cname].split(";") if d.strip()]
        assert list(dt.names) == fields, cname
        assert dt.itemsize == size, cname
