"""Synthetic VELO events (Allen raw-bank layout + Monte-Carlo truth).

Thin ctypes wrapper of ``libvelo_synth.so`` (built from ``velo_synth.cpp``).
The event model and the bank encoding are described in ``velo_synth.h``.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libvelo_synth.so")
DATA_DIR = os.path.join(os.path.dirname(_HERE), "data")


class SynthConfig(ctypes.Structure):
    _fields_ = [
        ("n_events", ctypes.c_uint),
        ("seed", ctypes.c_uint64),
        ("pileup", ctypes.c_int),
        ("mean_tracks", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("nb_mean", ctypes.c_double),
        ("nb_k", ctypes.c_double),
        ("high_occupancy_fraction", ctypes.c_double),
        ("high_nu", ctypes.c_double),
        ("eta_min", ctypes.c_double),
        ("eta_max", ctypes.c_double),
        ("backward_fraction", ctypes.c_double),
        ("pv_sigma_xy", ctypes.c_double),
        ("pv_sigma_z", ctypes.c_double),
        ("hit_efficiency", ctypes.c_double),
        ("noise_fraction", ctypes.c_double),
        ("dead_asic_fraction", ctypes.c_double),
        ("cluster_size_p", ctypes.c_double * 3),
        ("bank_version", ctypes.c_int),
        ("shuffle_sp", ctypes.c_int),
        ("retina", ctypes.c_int),
    ]


PARTICLE_DTYPE = np.dtype([
    ("event", "<u4"), ("pv", "<u4"), ("x0", "<f4"), ("y0", "<f4"), ("z0", "<f4"),
    ("tx", "<f4"), ("ty", "<f4"), ("eta", "<f4"), ("phi", "<f4"),
    ("n_hits", "<u4"), ("n_modules", "<u4"), ("first_module", "<u4"), ("last_module", "<u4"),
    ("p", "<f4"), ("pt", "<f4"),
])
HIT_DTYPE = np.dtype([
    ("event", "<u4"), ("particle", "<u4"), ("lhcb_id", "<u4"), ("sensor", "<u2"),
    ("n_pixels", "<u2"), ("x", "<f4"), ("y", "<f4"), ("z", "<f4"),
])

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(_LIB_PATH)
        lib.velo_synth_default_config.argtypes = [ctypes.POINTER(SynthConfig)]
        lib.velo_synth_generate.restype = ctypes.c_void_p
        lib.velo_synth_generate.argtypes = [ctypes.POINTER(SynthConfig), ctypes.c_char_p, ctypes.c_size_t,
                                            ctypes.c_uint]
        lib.velo_synth_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_size_t)]
        lib.velo_synth_free.argtypes = [ctypes.c_void_p]
        _lib = lib
    return _lib


def read_geometry() -> bytes:
    with open(os.path.join(DATA_DIR, "velo_geometry.bin"), "rb") as f:
        return f.read()


def read_beamline() -> np.ndarray:
    return np.fromfile(os.path.join(DATA_DIR, "beamline.bin"), dtype="<f4")


@dataclasses.dataclass
class Events:
    """A batch of events in Allen's layout: raw blocks, offsets, sizes and types."""
    n_events: int
    bank_version: int
    raw: np.ndarray          # uint8
    offsets: np.ndarray      # uint32 (n_events + 1)
    sizes: np.ndarray        # uint32 words
    types: np.ndarray        # uint32 words
    particles: np.ndarray | None = None
    particle_offsets: np.ndarray | None = None
    hits: np.ndarray | None = None
    hit_offsets: np.ndarray | None = None
    retina: bool = False     # banks are retina clusters (type 73) rather than super pixels
    mep: bool = False        # MEP layout (main/include/MEPTools.h:59-200) instead of Allen's

    @property
    def n_sp_words(self) -> int:
        sizes16 = self.sizes.view("<u2")
        start = int(self.sizes[0])
        n = int(sizes16[start:int(self.sizes[self.n_events])].astype(np.int64).sum()) // 4
        if self.bank_version in (2, 3):
            n -= int(self.sizes[self.n_events]) - start  # one count word per bank
        return n


def to_mep(ev: Events) -> Events:
    """The same events in the MEP layout that Allen's kernels read with mep_layout = true
    (main/include/MEPTools.h:59-200): one fragment per (bank, event), grouped by bank.

    - offsets (u32): [0] the number of banks, [1] the number of events, [2 + b] the source ID of
      bank b, [2 + n_banks (1 + e) + b] the byte offset of fragment (e, b) in the blocks;
    - blocks: for each bank, the fragments of every event in event order; a fragment is the bank
      without its source ID word (versions 2/3 keep their count word), padded to 4 bytes;
    - sizes (u32 + u16): [b] the position (in u16 units) of bank b's sizes, one per event;
    - types (u32 + u8): [b] the position (in bytes) of bank b's types, one per event.
    Every event must carry the same banks in the same order (as the fragments of one MEP do)."""
    if ev.mep:
        return ev
    n = ev.n_events
    s16 = ev.sizes.view("<u2")
    t8 = ev.types.view(np.uint8)
    per_event = []
    for e in range(n):
        base = int(ev.offsets[e])
        nb = int(ev.raw[base:base + 4].view("<u4")[0])
        boff = ev.raw[base + 4:base + 8 + 4 * nb].view("<u4").astype(np.int64)
        payload = base + 8 + 4 * nb
        sizes = s16[int(ev.sizes[e]):int(ev.sizes[e]) + nb].astype(np.int64)
        types = t8[int(ev.types[e]):int(ev.types[e]) + nb]
        src = [int(ev.raw[payload + o:payload + o + 4].view("<u4")[0]) for o in boff[:nb]]
        per_event.append((nb, boff, payload, sizes, types, src))
    nb = per_event[0][0] if n else 0
    sources = per_event[0][5] if n else []
    for e, (k, _, _, _, _, src) in enumerate(per_event):
        if k != nb or src != sources:
            raise ValueError(f"event {e} does not carry the first event's banks: no MEP layout")
    offsets = np.zeros(2 + nb * (1 + n), np.uint32)
    offsets[0], offsets[1] = nb, n
    offsets[2:2 + nb] = sources
    blocks = []
    pos = 0
    sizes_out = np.zeros(2 * nb + nb * n + (nb * n) % 2, np.uint16)
    types_out = np.zeros(4 * nb + nb * n + (-(4 * nb + nb * n)) % 4, np.uint8)
    sizes_out.view("<u4")[:nb] = 2 * nb + np.arange(nb) * n
    types_out.view("<u4")[:nb] = 4 * nb + np.arange(nb) * n
    for b in range(nb):
        for e, (_, boff, payload, sizes, types, _) in enumerate(per_event):
            start = payload + int(boff[b]) + 4
            end = payload + int(boff[b + 1]) if b + 1 < nb else payload + int(boff[nb])
            frag = ev.raw[start:end]
            pad = (-frag.size) % 4
            offsets[2 + nb * (1 + e) + b] = pos
            blocks.append(frag)
            if pad:
                blocks.append(np.zeros(pad, np.uint8))
            pos += frag.size + pad
            sizes_out[2 * nb + b * n + e] = sizes[b]
            types_out[4 * nb + b * n + e] = types[b]
    raw = np.concatenate(blocks) if blocks else np.zeros(4, np.uint8)
    return dataclasses.replace(ev, raw=np.ascontiguousarray(raw), offsets=offsets,
                               sizes=sizes_out.view("<u4").copy(), types=types_out.view("<u4").copy(), mep=True)


PRESETS = {
    # 200 events, Poisson(120) tracks from one PV, 3 % noise pixels.
    "single_pv_200": dict(n_events=200, pileup=0, mean_tracks=120.0),
    # 1000 events, nu ~ Poisson(7.6) collisions x NegBinomial(16, k=1.5) tracks.
    "pileup_1000": dict(n_events=1000, pileup=1, nu=7.6),
    # 20k events as single_pv_200 (one PV, Poisson(120) tracks, 3 % noise), of which 5 % are
    # high-occupancy events of nu = 20 pile-up collisions.
    "batched_20k": dict(n_events=20000, pileup=0, mean_tracks=120.0, high_occupancy_fraction=0.05, high_nu=20.0),
    # 10k events at nu ~ Poisson(15), 5 % noise, 1 % dead ASICs.
    "stress_10k": dict(n_events=10000, pileup=1, nu=15.0, noise_fraction=0.05, dead_asic_fraction=0.01),
}


def generate(n_threads: int = 8, keep_truth: bool = True, **overrides) -> Events:
    lib = _load()
    cfg = SynthConfig()
    lib.velo_synth_default_config(ctypes.byref(cfg))
    for k, v in overrides.items():
        if k == "cluster_size_p":
            for i in range(3):
                cfg.cluster_size_p[i] = v[i]
        else:
            setattr(cfg, k, v)
    geom = read_geometry()
    h = lib.velo_synth_generate(ctypes.byref(cfg), geom, len(geom), n_threads)
    if not h:
        raise ValueError("velo_synth_generate failed (bad geometry or bank version)")
    try:
        def get(name, dtype):
            p = ctypes.c_void_p()
            n = ctypes.c_size_t()
            if lib.velo_synth_get(h, name.encode(), ctypes.byref(p), ctypes.byref(n)):
                raise KeyError(name)
            if n.value == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ctypes.c_char * n.value).from_address(p.value)
            return np.frombuffer(bytes(buf), dtype=dtype).copy()

        ev = Events(
            n_events=cfg.n_events, bank_version=cfg.bank_version,
            raw=get("raw", np.uint8), offsets=get("offsets", np.uint32),
            sizes=get("sizes", np.uint32), types=get("types", np.uint32), retina=bool(cfg.retina))
        if keep_truth:
            ev.particles = get("particles", PARTICLE_DTYPE)
            ev.particle_offsets = get("particle_offsets", np.uint32)
            ev.hits = get("hits", HIT_DTYPE)
            ev.hit_offsets = get("hit_offsets", np.uint32)
        return ev
    finally:
        lib.velo_synth_free(h)


def preset(name: str, **overrides) -> Events:
    kw = dict(PRESETS[name])
    kw.update(overrides)
    return generate(**kw)
