"""getGPUImage's frame preparation on the device (cuda_bimp.cu:156-177) against the oracle:
the bilinear minimum-size upscale bit-exact for gray and BGR frames, the BGR -> gray
conversion of the upscaled frame, and detection on the top-left width x height of the
prepared frame against the FP64 oracle (the reference's CudaStuff keeps the original size)."""
import numpy as np
import pytest

import oracle_lib as O
from kpmatch import agreement, rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def _gray(bgr):
    b, g, r = (bgr[..., k].astype(np.uint32) for k in range(3))
    return ((1868 * b + 9617 * g + 4899 * r + (1 << 13)) >> 14).astype(np.uint8)


@pytest.mark.parametrize("w,h,cn", [(320, 240, 1), (200, 150, 3), (640, 240, 1), (349, 500, 3),
                                    (100, 61, 1)])
def test_upscale_bit_exact(B, w, h, cn):
    rng = np.random.default_rng(w + h)
    img = rng.integers(0, 256, (h, w) if cn == 1 else (h, w, cn), dtype=np.uint8)
    pw, ph, sf = B.min_size_shape(w, h)
    assert sf > 1
    got = B.resize_linear(img, pw, ph, sf, sf)
    want = O.resize_linear(img, pw, ph, sf, sf)
    assert np.array_equal(got, want)
    prep = B.get_gpu_image(img, grayscale=True)
    assert np.array_equal(prep, _gray(want) if cn == 3 else want)
    if cn == 3:
        assert np.array_equal(B.get_gpu_image(img, grayscale=False), want)


def test_no_upscale_is_a_copy(B):
    img = np.stack([O.synth_frame(640, 480, 10, 3 + c) for c in range(3)], axis=-1)
    assert np.array_equal(B.get_gpu_image(img, grayscale=False), img)
    assert np.array_equal(B.get_gpu_image(img), _gray(img))


@pytest.mark.parametrize("w,h", [(320, 240), (300, 200)])
def test_detect_on_cropped_upscale(B, w, h):
    img = O.synth_frame(w, h, 25, 44)
    prep = B.get_gpu_image(img)
    crop = np.ascontiguousarray(prep[:h, :w])
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    got, n = cs.detect(crop)
    orc = O.Oracle(w, h, O.BASE_BANK)
    kps, m, Ds, _ = orc.detect(crop, want_maps=True)
    assert max(rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i]) for i in range(len(cs.ks))) < 1e-4
    assert agreement(got, kps) >= 0.99
