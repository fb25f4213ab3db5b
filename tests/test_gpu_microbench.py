"""The measured device peaks behind the roofline figures (velo_microbench) are positive and no
higher than the published MI355X peaks allow (a measurement above them would mean the benchmark
does less work than it counts)."""
import pytest

from allen_amd import velo

# Published MI355X figures: 256 CUs x 64 lanes x 2.4 GHz non-packed FP32 lane operations,
# 256 CUs x 256 B per clock of LDS for 128-bit reads (CDNA4 doubled CDNA3's 128 B), 8 TB/s of HBM3E.
PUBLISHED = {
    "valu_lane_ops_per_s": 256 * 64 * 2.4e9,
    "lds_bytes_per_s": 256 * 256 * 2.4e9,
    "hbm_copy_bytes_per_s": 8e12,
}


@pytest.mark.gpu
@pytest.mark.parametrize("kind", sorted(PUBLISHED))
def test_microbench_within_published_peak(kind):
    value = velo.microbench(kind)
    print(f"{kind}: {value:.4g} ({value / PUBLISHED[kind]:.3f} of the published peak)", flush=True)
    assert 0.0 < value <= 1.05 * PUBLISHED[kind]
