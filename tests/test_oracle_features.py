"""The oracle's restatements of the keypoints' consumers, pinned without the reference:

- colour histogram (addColorFeatures_kernel, cuda_bimp.cuh:8-71): hand-placed pixels with
  known bins (the signed-char reading leaves bytes 0 and >= 128 uncounted), the window size
  p = (int)(log2(size)/2) at and next to powers of four, the window clipped at the frame, and
  a frame-free keypoint (0/0 -> NaN, as the reference's division gives);
- 2-NN matching: numpy brute force with a stable sort (ties to the lower index);
- the ratio test of the HIP library's host function (downloadMatches, bimp.cpp:208-240):
  known accept/reject cases around 0.6 (no GPU needed).
"""
import numpy as np

import bimp_amd as B
import oracle_lib as O


def _kp(points):
    """7 x n keypoint matrix from (x, y, size) triples."""
    m = np.zeros((7, len(points)), np.float32)
    for i, (x, y, s) in enumerate(points):
        m[0, i], m[1, i], m[4, i] = x, y, s
    return m


def test_color_histogram_known_bins():
    img = np.zeros((20, 20, 3), np.uint8)
    # keypoint (10.7, 10.2), size 4 -> p = 1: pixels x in {9, 10}, y in {9, 10}
    img[9:11, 9:11, 0] = [[1, 31], [32, 63]]      # B: bins 0, 0, 1, 1
    img[9:11, 9:11, 1] = [[64, 95], [96, 127]]    # G: bins 2, 2, 3, 3
    img[9:11, 9:11, 2] = [[0, 128], [200, 255]]   # R: counts nowhere (signed char <= 0)
    d = O.color_features(_kp([(10.7, 10.2, 4.0)]), img)[0]
    want = np.zeros(24)
    found = 4 * 3
    want[0] = want[1] = 2
    want[8 + 2] = want[8 + 3] = 2
    assert np.array_equal(d, (0.5 * want / found).astype(np.float32))
    assert np.all(d[4:8] == 0) and np.all(d[12:16] == 0) and np.all(d[16:] == 0)


def test_color_histogram_window_size():
    img = np.full((40, 40, 3), 50, np.uint8)  # every byte in bin 1
    for size, p in [(4.0, 1), (15.999998, 1), (16.0, 2), (63.999992, 2), (64.0, 3), (3.9, 0)]:
        d = O.color_features(_kp([(20.0, 20.0, size)]), img)[0]
        if p == 0:
            assert np.all(np.isnan(d))  # empty window: 0 / 0
            continue
        # every pixel counts once per channel: bin 1 = (2p)^2, found = 3 (2p)^2
        assert d[1] == np.float32(0.5 / 3) and d[9] == d[1] and d[17] == d[1], size
        assert np.count_nonzero(d) == 3


def test_color_histogram_clipped_window():
    img = np.zeros((10, 12, 3), np.uint8)
    img[0, 0] = (10, 40, 70)
    img[0, 1] = (10, 40, 70)
    img[1, 0] = (10, 40, 70)
    img[1, 1] = (10, 40, 70)
    # size 16 -> p = 2: window x, y in [-2, 2): only the 2 x 2 corner is inside
    d = O.color_features(_kp([(0.5, 0.9, 16.0)]), img)[0]
    assert d[0] == np.float32(0.5 * 4 / 12) and d[9] == d[0] and d[16 + 2] == d[0]
    # last column/row: the window [w-2, w+2) keeps 2 columns
    img2 = np.full((10, 12, 3), 5, np.uint8)
    d2 = O.color_features(_kp([(11.0, 9.0, 16.0)]), img2)[0]
    assert d2[0] == np.float32(0.5 / 3)
    # no pixel inside: NaN, as 0/0 in the reference
    assert np.all(np.isnan(O.color_features(_kp([(-10.0, -10.0, 16.0)]), img2)[0]))


def test_color_histogram_matches_numpy_restatement():
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (60, 80, 3), dtype=np.uint8)
    pts = [(rng.uniform(-3, 83), rng.uniform(-3, 63), rng.choice([4, 5.66, 8, 11.3, 16, 22.6, 45]))
           for _ in range(200)]
    got = O.color_features(_kp(pts), img)
    edges = np.arange(9) * (255.0 / 8)
    for i, (x, y, s) in enumerate(pts):
        p = int(np.floor(np.log2(np.float64(np.float32(s))) / 2))
        x0, y0 = int(np.float32(x)), int(np.float32(y))  # C casts truncate toward zero
        ys = [v for v in range(y0 - p, y0 + p) if 0 <= v < 60]
        xs = [v for v in range(x0 - p, x0 + p) if 0 <= v < 80]
        win = img[np.ix_(ys, xs)].reshape(-1, 3).view(np.int8).astype(np.float64)
        want = np.zeros(24)
        for ch in range(3):
            v = win[:, ch]
            for k in range(8):
                want[8 * ch + k] = np.count_nonzero((v > edges[k]) & (v < edges[k + 1]))
        found = 3 * len(ys) * len(xs)
        with np.errstate(invalid="ignore", divide="ignore"):
            want = (0.5 * want / found).astype(np.float32)
        assert np.array_equal(got[i], want, equal_nan=True), i


def test_knn2_matches_numpy_brute_force():
    rng = np.random.default_rng(1)
    q = rng.standard_normal((50, 24)).astype(np.float32)
    t = rng.standard_normal((70, 24)).astype(np.float32)
    t[10] = t[3]  # a tie: the lower index comes first
    q[0] = t[3]
    idx, dist = O.knn2(q, t)
    d = np.sqrt(((q[:, None, :].astype(np.float64) - t[None, :, :]) ** 2).sum(-1))
    order = np.argsort(d, axis=1, kind="stable")[:, :2]
    assert np.array_equal(idx, order)
    assert np.allclose(dist, np.take_along_axis(d, order, 1), rtol=1e-6)
    assert idx[0, 0] == 3 and idx[0, 1] == 10 and dist[0, 0] == 0
    # fewer than two train rows
    i1, d1 = O.knn2(q[:2], t[:1])
    assert np.all(i1[:, 1] == -1) and np.all(np.isinf(d1[:, 1]))


def test_ratio_matches_known_answers():
    idx = np.array([[4, 7], [1, 2], [9, 3], [5, -1], [-1, -1]], np.int32)
    dist = np.array([[1.0, 2.0], [0.6, 1.0], [0.59, 1.0], [0.1, np.inf], [np.inf, np.inf]],
                    np.float32)
    m = B.ratio_matches(idx, dist)
    # 1 < 1.2 keeps row 0; 0.6f < 0.6 fails (0.6f > 0.6 in double); 0.59 keeps; inf keeps
    assert list(m["query_idx"]) == [0, 2, 3]
    assert list(m["train_idx"]) == [4, 9, 5]
    assert np.array_equal(m["distance"], np.float32([1.0, 0.59, 0.1]))
    assert np.all(m["img_idx"] == 0)


# ------------------------------------------------------------------ frame preparation
def _shape_rule(w, h):
    """getGPUImage (cuda_bimp.cu:159-168) and cv::resize's output size, from the text."""
    if h < 350 and w >= 350:
        sf = np.float32(350 / h)
    elif w < 350:
        sf = np.float32(350 / w)
    else:
        return w, h
    return int(np.rint(w * np.float64(sf))), int(np.rint(h * np.float64(sf)))


def test_min_size_shape_rule():
    for w in list(range(1, 800, 7)) + [349, 350, 351]:
        for h in list(range(1, 800, 11)) + [349, 350, 351]:
            ow, oh, f = B.min_size_shape(w, h)
            assert (ow, oh) == _shape_rule(w, h), (w, h)
            assert (f == 1.0) == ((ow, oh) == (w, h))


def test_resize_known_answers():
    # constant frames stay constant
    c = np.full((30, 40), 137, np.uint8)
    assert np.all(O.resize_linear(c, 77, 59, 77 / 40, 59 / 30) == 137)
    # factor 2: output x maps to source (x + 0.5) / 2 - 0.5, so pairs of outputs sit a quarter
    # pixel either side of a source sample (weights 1536/512), clamped at the ends
    row = np.array([[0, 128, 255, 64]], np.uint8)
    out = O.resize_linear(np.repeat(row, 3, 0), 8, 6, 2.0, 2.0)[0]
    want = [0, 32, 96, 160, 223, 207, 112, 64]
    assert list(out) == want
    # channels are independent
    rgb = np.stack([np.repeat(row, 3, 0), np.repeat(row, 3, 0) // 2, np.full((3, 4), 9, np.uint8)], -1)
    o3 = O.resize_linear(rgb, 8, 6, 2.0, 2.0)
    assert list(o3[0, :, 0]) == want and np.all(o3[..., 2] == 9)


def test_resize_close_to_float_bilinear():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (120, 160), dtype=np.uint8)
    sf = np.float32(350 / 160)
    dw, dh = _shape_rule(160, 120)
    got = O.resize_linear(img, dw, dh, float(sf), float(sf)).astype(np.int32)
    inv = 1.0 / np.float64(sf)

    def coords(n_out, n_in):
        f = (np.arange(n_out) + 0.5) * inv - 0.5
        s = np.floor(f).astype(int)
        fr = f - s
        fr[s < 0] = 0
        s[s < 0] = 0
        fr[s >= n_in - 1] = 0
        s[s >= n_in - 1] = n_in - 1
        return s, np.minimum(s + 1, n_in - 1), fr

    x0, x1, fx = coords(dw, 160)
    y0, y1, fy = coords(dh, 120)
    im = img.astype(np.float64)
    top = im[y0][:, x0] * (1 - fx) + im[y0][:, x1] * fx
    bot = im[y1][:, x0] * (1 - fx) + im[y1][:, x1] * fx
    ref = top * (1 - fy)[:, None] + bot * fy[:, None]
    assert np.abs(got - ref).max() <= 1.0


# ------------------------------------------------------------------ SURF descriptors
def test_integral_image_and_haar_boxes():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    ii = O.integral_image(img)
    assert np.array_equal(ii[1:, 1:], img.astype(np.int64).cumsum(0).cumsum(1))
    assert np.all(ii[0] == 0) and np.all(ii[:, 0] == 0)


def _edge(h, w, vertical):
    img = np.full((h, w), 40, np.uint8)
    if vertical:
        img[:, w // 2:] = 200   # dark left, bright right
    else:
        img[h // 2:, :] = 200   # dark top, bright bottom
    return img


def test_surf_step_edges():
    h, w = 120, 120
    for vertical in (True, False):
        img = _edge(h, w, vertical)
        kp = _kp([(60.0, 60.0, 16.0)])
        d, a = O.surf_descriptors(img, kp)
        # every Haar response points across the edge: +x (angle 0) or +y (angle pi/2)
        assert a[0] == (0.0 if vertical else np.pi / 2), a
        assert abs(np.linalg.norm(d[0].astype(np.float64)) - 1) < 1e-6
        du, _ = O.surf_descriptors(img, kp, upright=True)
        f = du[0].reshape(16, 4)
        if vertical:
            assert np.all(f[:, 1] == 0) and np.all(f[:, 3] == 0)      # no dy
            assert np.all(f[:, 0] >= 0) and np.array_equal(f[:, 0], f[:, 2])
        else:
            assert np.all(f[:, 0] == 0) and np.all(f[:, 2] == 0)      # no dx
            assert np.all(f[:, 1] >= 0) and np.array_equal(f[:, 1], f[:, 3])
        # in its own frame the oriented descriptor of either edge puts the response on the
        # first axis: the same descriptor for both edges
        fo = d[0].reshape(16, 4)
        assert np.allclose(fo[:, 1], 0, atol=1e-6) and np.allclose(fo[:, 3], 0, atol=1e-6)
        if vertical:
            ref = fo.copy()
        else:
            assert np.allclose(fo, ref, atol=1e-6)


def test_surf_contrast_invariance_and_norm():
    rng = np.random.default_rng(8)
    img = rng.integers(0, 128, (90, 110), dtype=np.uint8)
    pts = [(rng.uniform(5, 105), rng.uniform(5, 85), rng.choice([4.0, 5.66, 8.0, 11.3, 16.0]))
           for _ in range(40)]
    kp = _kp(pts)
    d1, a1 = O.surf_descriptors(img, kp)
    d2, a2 = O.surf_descriptors(img * 2, kp)
    assert np.allclose(a1, a2, atol=1e-12)
    assert np.allclose(d1, d2, atol=1e-6)
    assert np.allclose(np.linalg.norm(d1.astype(np.float64), axis=1), 1, atol=1e-6)
