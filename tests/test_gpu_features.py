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


def _wrap(a):
    return (a + np.pi) % (2 * np.pi) - np.pi


@pytest.mark.parametrize("upright", [True, False])
def test_surf_matches_oracle(B, upright):
    """SURF descriptors of the detector's keypoints against the oracle's FP64 restatement.
    Upright, the sample positions are identical and only the sums' rounding differs (every
    keypoint within 1e-5). Oriented, the angle comes from FP32 against FP64 window sums, so a
    sample on a window or rounding boundary can move; >= 97% of keypoints agree within 1e-3."""
    w, h = 640, 480
    gray = O.synth_frame(w, h, 40, 0)
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    mat, n = cs.detect_matrix(gray)
    gd, ga = B.surf_descriptors(gray, mat, upright=upright)
    od, oa = O.surf_descriptors(gray, mat, upright=upright)
    assert np.allclose(np.linalg.norm(gd.astype(np.float64), axis=1), 1, atol=1e-5)
    dmax = np.abs(gd.astype(np.float64) - od).max(axis=1)
    amax = np.abs(_wrap(ga.astype(np.float64) - oa))
    if upright:
        assert np.all(ga == 0) and dmax.max() <= 1e-5
        return
    ok = (dmax <= 1e-3) & (amax <= 1e-3)
    print(f"{n} keypoints: {ok.mean():.4f} within 1e-3 (descriptor max diff median "
          f"{np.median(dmax):.1e}, angle diff median {np.median(amax):.1e})", flush=True)
    assert ok.mean() >= 0.97


def test_surf_matching_finds_shift(B):
    """Descriptors + 2-NN ratio matching between a frame and the same scene shifted by
    (+7, +4) pixels: most ratio-test matches pair a keypoint with its shifted twin."""
    w, h = 640, 480
    big = O.synth_frame(w + 16, h + 16, 40, 3)
    a = np.ascontiguousarray(big[8:8 + h, 8:8 + w])
    b = np.ascontiguousarray(big[4:4 + h, 1:1 + w])  # a's content appears at (+7, +4) in b
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    ma, na = cs.detect_matrix(a)
    mb, nb = cs.detect_matrix(b)
    da, _ = B.surf_descriptors(a, ma)
    db, _ = B.surf_descriptors(b, mb)
    idx, dist = B.match_knn2(da, db)
    m = B.ratio_matches(idx, dist)
    qa, tb = m["query_idx"], m["train_idx"]
    dx = mb[0, tb] - ma[0, qa]
    dy = mb[1, tb] - ma[1, qa]
    good = (np.abs(dx - 7) <= 1) & (np.abs(dy - 4) <= 1)
    print(f"{na} / {nb} keypoints, {len(m)} ratio matches, {good.mean():.3f} at the shift",
          flush=True)
    assert len(m) > 100 and good.mean() >= 0.8
