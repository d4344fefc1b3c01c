"""GPU parity of the precision form (BF16 / FP16 direct convolution on the matrix cores).

Two bars per direct-form scale:
  - against the oracle's restatement of the same arithmetic (same rounded taps, FP64 sums):
    within 5e-5 relative (the library accumulates in FP32);
  - against the FP64 FFT oracle (the reference's filter): within 1e-2 relative per map.
End to end, keypoints agree with the FP64 oracle's at >= 99% on the 640x480 frame.
Scales above direct_max_lambda keep the FP32 FFT form and its 1e-4 bar.
"""
import os

import numpy as np
import pytest

import oracle_lib as O
from kpmatch import agreement, rel_err

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def _opts(B, precision, nsig=0.0, max_lambda=8.0):
    o = B.default_options()
    o.precision = precision
    o.direct_sigmas = nsig
    o.direct_max_lambda = max_lambda
    return o


CASES = [("bf16", 0.0), ("fp16", 0.0), ("bf16", 6.0), ("fp16", 8.0)]


def _prec(B, name):
    return B.PRECISION_FP16 if name == "fp16" else B.PRECISION_BF16


@pytest.mark.parametrize("fmt,nsig", CASES)
@pytest.mark.parametrize("w,h", [(160, 120), (200, 150)])
def test_direct_maps_match_restatement(B, w, h, fmt, nsig):
    lam = O.BASE_BANK
    cs = B.CudaStuff(w, h, lam, _opts(B, _prec(B, fmt), nsig))
    orc = O.Oracle(w, h, lam)
    img = O.synth_frame(w, h, 12, 21)
    orc.set_image(img)
    cs.filter(img)
    ndirect = 0
    for i, si in enumerate(cs.ks):
        direct = si.lambda2 <= 8.0
        assert si.precision == (_prec(B, fmt) if direct else B.PRECISION_FP32), i
        if direct:
            ndirect += 1
            ref = orc.direct_filter(i, nsig, fmt == "fp16")
            tol = 5e-5
        else:
            ref = orc.filter(i)
            tol = 1e-4
        for o in range(8):
            got = cs.get_map(i, B.MAP_COMPLEX, o)
            e = rel_err(got, ref[o, :si.map_h, :si.map_w])
            print(f"{fmt} nsig {nsig} scale {i} ext {si.tap_extent} o {o} rel_err {e:.2e}",
                  flush=True)
            assert e < tol, (i, o, e)
    assert ndirect == 5


@pytest.mark.parametrize("fmt,nsig", [("bf16", 0.0), ("fp16", 0.0), ("bf16", 8.0)])
def test_direct_maps_within_1e2_of_fp64(B, fmt, nsig):
    w, h = 320, 240
    lam = O.BASE_BANK[:5]
    cs = B.CudaStuff(w, h, lam, _opts(B, _prec(B, fmt), nsig))
    orc = O.Oracle(w, h, lam)
    img = O.synth_frame(w, h, 25, 5)
    orc.set_image(img)
    cs.filter(img)
    for i, si in enumerate(cs.ks):
        ref = orc.filter(i)
        worst = max(rel_err(cs.get_map(i, B.MAP_COMPLEX, o), ref[o, :si.map_h, :si.map_w])
                    for o in range(8))
        print(f"{fmt} nsig {nsig} scale {i} lambda2 {si.lambda2:.3f} max rel err {worst:.2e}",
              flush=True)
        assert worst <= 1e-2


@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_direct_keypoints_agree_on_640x480(B, fmt):
    """The 640x480 frame (40 shapes, seed 0) against the FP64 oracle's keypoints."""
    w, h = 640, 480
    cs = B.CudaStuff(w, h, O.BASE_BANK, _opts(B, _prec(B, fmt)))
    img = O.synth_frame(w, h, 40, 0)
    want = np.load(os.path.join(GOLDEN, "cfg0_keypoints.npy"), allow_pickle=False)
    kps = np.ascontiguousarray(want).view(O.KP_DTYPE).reshape(-1)  # oracle keypoint records
    got, gn = cs.detect(img)
    a = agreement(got, kps)
    print(f"{fmt}: keypoints gpu {gn} oracle {len(kps)} agreement {a:.4f}", flush=True)
    assert a >= 0.99


def test_all_scales_direct_and_repeatable(B):
    """direct_max_lambda above the bank: every scale direct; two runs bit-identical."""
    w, h = 200, 150
    cs = B.CudaStuff(w, h, O.BASE_BANK, _opts(B, B.PRECISION_BF16, 0.0, 1e9))
    assert all(si.precision == B.PRECISION_BF16 for si in cs.ks)
    img = O.synth_frame(w, h, 15, 8)
    k1, n1 = cs.detect(img)
    m1 = [cs.get_map(i, B.MAP_DOUBLE).copy() for i in range(len(cs.ks))]
    k2, n2 = cs.detect(img)
    assert n1 == n2 and np.array_equal(k1, k2)
    for i in range(len(cs.ks)):
        assert np.array_equal(m1[i], cs.get_map(i, B.MAP_DOUBLE))


def test_direct_wraps_like_fft_on_narrow_frames(B):
    """Kernel extent beyond the transform size: the direct form wraps as the FFT does."""
    w, h = 40, 24  # FFT 64 x 32, kW up to 47
    lam = O.BASE_BANK[:4]
    cs = B.CudaStuff(w, h, lam, _opts(B, B.PRECISION_FP16))
    orc = O.Oracle(w, h, lam)
    img = O.synth_frame(w, h, 3, 2)
    orc.set_image(img)
    cs.filter(img)
    for i, si in enumerate(cs.ks):
        ref = orc.direct_filter(i, 0.0, True)
        for o in range(8):
            e = rel_err(cs.get_map(i, B.MAP_COMPLEX, o), ref[o, :si.map_h, :si.map_w])
            assert e < 5e-5, (i, o, e)
