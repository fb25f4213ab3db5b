"""MDF ingestion: the VELO raw banks of an MDF file in Allen's event layout.

Host side of `velo_mdf_read` (allen_amd/csrc/velo_mdf.cpp). The result is an `allen_amd.synth.Events`
without truth, ready for `velo.DeviceEvents`:

    from allen_amd import mdf, velo
    ev = mdf.read("minbias.mdf", mdf.RETINA)      # retina cluster banks (type 73)
    seq.run(velo.DeviceEvents(ev))
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import velo
from .synth import Events

VP = 63        # VELO super-pixel banks (LHCb::RawBank::VP)
RETINA = 73    # VELO retina cluster banks (LHCb::RawBank::VPRetinaCluster)


def read(path: str, bank_type: int = VP, max_events: int = 0) -> Events:
    """Banks of `bank_type` of the first `max_events` records (0: all) of an uncompressed MDF file."""
    L = velo.lib()
    h = ctypes.c_void_p()
    velo.check(L.velo_mdf_read(str(path).encode(), int(bank_type), int(max_events), ctypes.byref(h)), "velo_mdf_read")
    try:
        n, ver, nbytes = ctypes.c_uint(), ctypes.c_int(), ctypes.c_size_t()
        velo.check(L.velo_mdf_events_info(h, ctypes.byref(n), ctypes.byref(ver), ctypes.byref(nbytes)), "info")

        def get(name):
            p, b = ctypes.c_void_p(), ctypes.c_size_t()
            velo.check(L.velo_mdf_events_buffer(h, name.encode(), ctypes.byref(p), ctypes.byref(b)), name)
            if b.value == 0:
                return np.zeros(0, np.uint32)
            return np.frombuffer(ctypes.string_at(p.value, b.value), dtype=np.uint32).copy()

        raw = get("raw").view(np.uint8)
        return Events(n_events=n.value, bank_version=ver.value, raw=raw, offsets=get("offsets"), sizes=get("sizes"),
                      types=get("types"), retina=bank_type == RETINA)
    finally:
        L.velo_mdf_events_free(h)
