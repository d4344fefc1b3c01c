"""GPU parity of the keypoints' consumers against the oracle.

- Colour histogram on the keypoints of a colour frame (its BGR->gray image through the
  detector): bit-exact against the oracle's restatement of addColorFeatures_kernel.
- 2-NN matching: train indices equal to the oracle's (FP64 sums) wherever its two best
  distances are not within 1e-5 of each other; distances within 1e-5 relative; the ratio test
  on the device's result keeps the oracle's matches.
"""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


@pytest.mark.parametrize("w,h", [(640, 480), (1920, 1080)])
def test_color_features_bit_exact_on_detected_keypoints(B, w, h):
    bgr = np.stack([O.synth_frame(w, h, 60, 70 + c) for c in range(3)], axis=-1)
    gray = B.bgr_to_gray(bgr)
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    mat, n = cs.detect_matrix(gray)
    assert n > 100
    got = B.color_features(mat, bgr)
    want = O.color_features(mat, bgr)
    assert got.shape == (mat.shape[1], 24)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.all(np.isfinite(got))


@pytest.mark.parametrize("nq,nt,dim", [(1000, 1500, 152), (300, 70, 64), (65, 1, 24)])
def test_knn2_matches_oracle(B, nq, nt, dim):
    rng = np.random.default_rng(nq + nt)
    t = rng.standard_normal((nt, dim)).astype(np.float32)
    q = (t[rng.integers(0, nt, nq)] + 0.3 * rng.standard_normal((nq, dim))).astype(np.float32)
    gi, gd = B.match_knn2(q, t)
    oi, od = O.knn2(q, t)
    fin = np.isfinite(od)
    assert np.array_equal(np.isfinite(gd), fin)
    assert np.allclose(gd[fin], od[fin], rtol=1e-5, atol=0)
    clear = ~fin[:, 1] | (np.abs(od[:, 1] - od[:, 0]) > 1e-5 * od[:, 1])
    assert np.array_equal(gi[clear], oi[clear])
    gm = B.ratio_matches(gi, gd)
    om = B.ratio_matches(oi, od)
    print(f"{nq}x{nt}x{dim}: {len(gm)} ratio matches (oracle {len(om)})", flush=True)
    assert len(gm) == len(om) and np.array_equal(gm["train_idx"], om["train_idx"])
