"""The transform micro-benchmark behind the VALU-or-MFMA decision for the
group filter's DCT: both forms run, and a forward + inverse round trip
returns the input to within FP32 rounding (an orthonormal transform on
0-255 data: errors of order 1e-4; a wrong table, layout or exchange would
give errors of the data's size)."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [7, 8])
@pytest.mark.parametrize("mode", [0, 1], ids=["valu", "mfma"])
def test_round_trip(K, mode):
    import vbm3d_amd.api as A
    ms, rate, err = A.bench_transforms(K, mode, npatch=4099, reps=3)
    assert ms > 0 and rate > 0
    assert err < 2e-3, err


def test_bad_arguments():
    import vbm3d_amd.api as A
    with pytest.raises(A.VBM3DError):
        A.bench_transforms(6, 0, 100, 1)
    with pytest.raises(A.VBM3DError):
        A.bench_transforms(7, 2, 100, 1)
