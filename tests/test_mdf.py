"""MDF ingestion (velo_mdf_read): real files, written files and malformed records (no GPU needed)."""
import os
import struct

import numpy as np
import pytest

from allen_amd import mdf, synth, velo

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REFERENCE_MDF = "/root/reference/input/minbias/mdf/MiniBrunel_2018_MinBias_FTv4_DIGI_retinacluster_v1.mdf"


def _banks(ev, e):
    """(type, sourceID, payload bytes) of every bank of event e in Allen's layout."""
    base = int(ev.offsets[e])
    n = int(ev.raw[base:base + 4].view("<u4")[0])
    off = ev.raw[base + 4:base + 4 + 4 * (n + 1)].view("<u4")
    payload = base + 4 + 4 * (n + 1)
    s16 = ev.sizes.view("<u2")[int(ev.sizes[e]):int(ev.sizes[e]) + n]
    t8 = ev.types.view(np.uint8)[int(ev.types[e]):int(ev.types[e]) + n]
    out = []
    for b in range(n):
        p = payload + int(off[b])
        out.append((int(t8[b]), int(ev.raw[p:p + 4].view("<u4")[0]), ev.raw[p + 4:p + 4 + int(s16[b])].tobytes()))
    return out


def write_mdf(path, records, compression=0, magic=0xCBCB):
    """records: per event a list of (type, version, sourceID, payload bytes); header version 3."""
    with open(path, "wb") as f:
        for banks in records:
            body = b""
            for btype, version, source, payload in banks:
                body += struct.pack("<HHBBH", magic, 8 + len(payload), btype, version, source) + payload
                body += b"\0" * ((-len(payload)) % 4)
            size = 48 + len(body)
            f.write(struct.pack("<4I4B", size, size, size, 0, compression, (3 << 4) | 7, 1, 0) + b"\0" * 28 + body)


@pytest.mark.skipif(not os.path.exists(REFERENCE_MDF), reason="the reference's MDF file is not on this machine")
@pytest.mark.parametrize("bank_type,fixture", [(mdf.VP, "mdf_minbias_vp.npz"), (mdf.RETINA, "mdf_minbias_retina.npz")])
def test_reader_reproduces_the_fixtures(bank_type, fixture):
    """The bundled MDF file read natively gives the committed fixtures (made by the independent
    Python parser in tests/golden) byte for byte."""
    ev = mdf.read(REFERENCE_MDF, bank_type)
    z = np.load(os.path.join(GOLDEN, fixture))
    assert ev.n_events == int(z["n_events"]) and ev.bank_version == int(z["version"])
    for k in ("raw", "offsets", "sizes", "types"):
        np.testing.assert_array_equal(getattr(ev, k), z[k])
    assert ev.retina == (bank_type == mdf.RETINA)


@pytest.mark.parametrize("retina", [0, 1])
def test_round_trip_through_a_written_file(retina, tmp_path):
    """Synthetic events written as MDF records, with their banks in reverse sourceID order and
    banks of other types mixed in, read back into the generator's own layout."""
    ev = synth.generate(n_events=5, pileup=1, nu=7.6, seed=71, retina=retina, keep_truth=False)
    version = ev.bank_version
    records = []
    for e in range(ev.n_events):
        banks = [(t, version, s & 0xFFFF, p) for t, s, p in _banks(ev, e)][::-1]
        banks.insert(3, (9, 1, 5, b"\1\2\3\4\5"))     # another detector's bank
        banks.append((16, 0, 0, b""))                # an empty bank
        records.append(banks)
    path = tmp_path / "events.mdf"
    write_mdf(path, records)
    got = mdf.read(str(path), mdf.RETINA if retina else mdf.VP)
    assert got.n_events == ev.n_events and got.bank_version == version
    for k in ("raw", "offsets", "sizes", "types"):
        np.testing.assert_array_equal(getattr(got, k), getattr(ev, k))
    assert mdf.read(str(path), mdf.VP if retina else mdf.RETINA, max_events=2).raw.size == 2 * 4 * 2


def test_malformed_files(tmp_path):
    payload = struct.pack("<2I", 1, 0x12345678)
    cases = {
        "compressed": (dict(compression=1), velo.VeloError, "compressed MDF record"),
        "bad_magic": (dict(magic=0xABCD), velo.VeloError, "well-formed MDF"),
    }
    for name, (kw, exc, text) in cases.items():
        path = tmp_path / f"{name}.mdf"
        write_mdf(path, [[(63, 2, 0, payload)]], **kw)
        with pytest.raises(exc, match=text):
            mdf.read(str(path))
    mixed = tmp_path / "mixed.mdf"
    write_mdf(mixed, [[(63, 2, 0, payload)], [(63, 4, 0, payload)]])
    with pytest.raises(velo.VeloError, match="version"):
        mdf.read(str(mixed))
    with pytest.raises(velo.VeloError, match="well-formed MDF"):
        mdf.read(str(tmp_path / "missing.mdf"))
