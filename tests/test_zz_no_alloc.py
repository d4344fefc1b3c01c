"""bimp_detect_u8_noalloc (keypoints staged in the library's static device array, no device
allocation per call) against bimp_detect_u8 on the same contexts and frames: bit-identical
records and counts on two calls in a row, the second with a smaller frame, for host and device
frames; deterministic truncation at a small cap; and the documented refusals."""
import numpy as np
import pytest

import oracle_lib as O
from kpmatch import bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def test_noalloc_equals_detect_two_calls(B):
    big = B.CudaStuff(640, 480, O.BASE_BANK)
    small = B.CudaStuff(320, 240, O.BASE_BANK)
    f_big = O.synth_frame(640, 480, 40, 0)
    f_small = O.synth_frame(320, 240, 25, 5)
    # two calls in a row, the second with a smaller input
    g1, n1 = big.detect_noalloc(f_big)
    g2, n2 = small.detect_noalloc(f_small)
    w1, m1 = big.detect(f_big)
    w2, m2 = small.detect(f_small)
    print(f"keypoints {n1} / {n2}", flush=True)
    assert n1 == m1 and n2 == m2 and n2 < n1
    assert np.array_equal(bits(g1), bits(w1))
    assert np.array_equal(bits(g2), bits(w2))


def test_noalloc_device_frame_and_cap(B):
    w, h = 320, 240
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = O.synth_frame(w, h, 30, 9)
    pitch = 384
    padded = np.zeros((h, pitch), np.uint8)
    padded[:, :w] = img
    d = B.DeviceBuffer(padded.nbytes)
    d.upload(padded)
    want, m = cs.detect(img)
    got, n = cs.detect_noalloc(d, pitch=pitch)
    assert n == m and np.array_equal(bits(got), bits(want))
    cap = m // 3
    part, n3 = cs.detect_noalloc(img, cap=cap)  # BIMP_W_OVERFLOW is a warning
    assert n3 == m and np.array_equal(bits(part), bits(want[:cap]))


def test_noalloc_refusals(B):
    cs = B.CudaStuff(160, 120, O.BASE_BANK)
    img = O.synth_frame(160, 120, 10, 2)
    with pytest.raises(B.BimpError) as e:
        cs.detect_noalloc(img, cap=65537)
    assert e.value.status == B.E_ARG
    cs.set_flags(B.FLAG_PROFILE)
    with pytest.raises(B.BimpError) as e:
        cs.detect_noalloc(img)
    assert e.value.status == B.E_UNSUPPORTED
    cs.set_flags(0)
    got, n = cs.detect_noalloc(img, cap=65536)
    want, m = cs.detect(img, cap=65536)
    assert n == m and np.array_equal(bits(got), bits(want))
