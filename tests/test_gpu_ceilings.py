"""The measured FP64 ceilings that bench.py sets beside the published FP64
peak: both loops run, and neither exceeds the 78.6 TFLOP/s the device is
rated at (a timing or flop-count slip in the probe would)."""
import pytest

pytestmark = pytest.mark.gpu


def test_fp64_ceilings_are_plausible():
    from redondito_amd import sphp as S
    mfma, valu = S.fp64_ceilings()
    print(f"FP64 ceilings: MFMA {mfma:.1f} TFLOP/s, VALU {valu:.1f} TFLOP/s")
    for v in (mfma, valu):
        assert 5.0 < v < 1.05 * 78.6


def test_fp64_mfma4x4_ceiling_is_plausible():
    from redondito_amd import sphp as S
    t = S.fp64_mfma4x4_ceiling()
    print(f"FP64 4x4x4 MFMA ceiling: {t:.1f} TFLOP/s")
    assert 5.0 < t < 1.05 * 78.6
