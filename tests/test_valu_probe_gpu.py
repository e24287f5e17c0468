"""The measured 32-bit integer VALU rate (fg_valu_probe) against the peak the
walk's roofline fraction is quoted against: CUs x clock x 128 lanes (two
wave64 VALU instructions per SIMD every 4 cycles on gfx950)."""
import pytest

import flipgraphgpu_amd as fg


@pytest.mark.gpu
def test_valu_probe_reaches_the_nominal_peak():
    info = fg.device_info()
    nominal = info["cu"] * info["clock_khz"] * 1e3 * 128
    ops, ms = fg.valu_probe()
    assert ms > 0
    # above what one VALU instruction per SIMD every 4 cycles allows (64
    # lanes), at most the peak at the reported clock
    assert 0.6 * nominal < ops < 1.02 * nominal, (ops, nominal)


@pytest.mark.gpu
def test_valu_probe_rejects_bad_arguments():
    with pytest.raises(fg.FgError):
        fg.valu_probe(0, 10)
    with pytest.raises(fg.FgError):
        fg.valu_probe(16, 0)
