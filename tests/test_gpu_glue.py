"""GPU test of getGPUImage's colour step (bimp_bgr_to_gray, cuda_bimp.cu:156-179).

OpenCV is not installed, so parity with cv::cuda::cvtColor is pinned by what its Q14 weights
fix: grey pixels are unchanged (the weights sum to 2^14), and each pure channel maps to
floor(255 * w / 2^14 + 1/2) with w = 1868 (B), 9617 (G), 4899 (R)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def test_grey_pixels_unchanged(B):
    v = np.arange(256, dtype=np.uint8).reshape(16, 16)
    bgr = np.repeat(v[:, :, None], 3, axis=2)
    assert np.array_equal(B.bgr_to_gray(bgr), v)


def test_channel_order_and_rounding(B):
    bgr = np.zeros((1, 4, 3), np.uint8)
    bgr[0, 0] = (255, 0, 0)      # blue
    bgr[0, 1] = (0, 255, 0)      # green
    bgr[0, 2] = (0, 0, 255)      # red
    bgr[0, 3] = (10, 20, 30)
    g = B.bgr_to_gray(bgr)[0]
    assert list(g[:3]) == [29, 150, 76]
    assert g[3] == (1868 * 10 + 9617 * 20 + 4899 * 30 + 8192) >> 14


@pytest.mark.parametrize("w,h", [(160, 120), (320, 240)])
def test_cpu_keypoint_mode_matches_oracle(B, w, h):
    """bimp_keypoints_cpu_mode equals the oracle's restatement of the reference's CPU mode on
    the same D and S maps, bit for bit."""
    import oracle_lib as O
    from kpmatch import bits
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = O.synth_frame(w, h, 20, 6)
    cs.detect(img)
    got, n = cs.keypoints_cpu_mode()
    orc = O.Oracle(w, h, O.BASE_BANK)
    want, m = [], 0
    for i in range(len(cs.ks)):
        k, c = orc.nms_cpu_mode(i, cs.get_map(i, B.MAP_DOUBLE), cs.get_map(i, B.MAP_SINGLE))
        want.append(k)
        m += c
    want = np.concatenate(want)
    assert n == m and n > 0
    assert np.array_equal(bits(got), bits(want))


def test_pitched_host_frame_equals_contiguous(B):
    """A host frame with a row pitch wider than its width (staged through pinned memory)
    gives the keypoints of the same frame passed contiguously."""
    import ctypes as C

    import oracle_lib as O
    w, h = 200, 150
    img = O.synth_frame(w, h, 12, 7)
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    want, wn = cs.detect(img)
    padded = np.full((h, w + 56), 255, np.uint8)
    padded[:, :w] = img
    for _ in range(2):  # the second call reuses the staging buffer
        out = np.zeros(B.KP_MAX_POINTS, B.KEYPOINT_DTYPE)
        n = C.c_int(0)
        st = B.lib.bimp_detect_u8(cs.ctx, B._ptr(padded), w + 56, 0, B._ptr(out), B.KP_MAX_POINTS,
                                  0, C.byref(n), None)
        assert st >= 0
        assert n.value == wn
        assert np.array_equal(out[:wn].view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("w,h,form", [(640, 480, "fp32"), (1920, 1080, "fp32"),
                                      (1920, 1080, "bf16")])
def test_two_streams_equal_one_stream(B, w, h, form):
    """Work above pyramid level 0 on the context's second stream (FFT groups; in the
    precision form, the level's own direct launch) gives the maps and keypoints of the same
    frame run on one stream (one direct launch for every scale), bit for bit: with the flag
    set at creation (no second stream made) and switched on and off at run time, on the null
    stream and on a created one."""
    import oracle_lib as O
    from kpmatch import bits
    img = O.synth_frame(w, h, 60, 8)

    def options(flags):
        opt = B.default_options()
        opt.flags = flags
        if form == "bf16":
            opt.precision = B.PRECISION_BF16
        return opt

    one = B.CudaStuff(w, h, O.BASE_BANK, options(B.FLAG_SINGLE_STREAM))
    want, wn = one.detect(img, cap=1 << 16)
    wmaps = [(one.get_map(i, B.MAP_DOUBLE), one.get_map(i, B.MAP_SINGLE)) for i in range(len(one.ks))]
    assert wn > 0
    two = B.CudaStuff(w, h, O.BASE_BANK, options(0))
    s = B.Stream()
    for flags, stream in [(0, None), (0, s.handle), (B.FLAG_SINGLE_STREAM, s.handle), (0, s.handle)]:
        two.set_flags(flags)
        got, n = two.detect(img, cap=1 << 16, stream=stream)
        assert n == wn
        assert np.array_equal(bits(got), bits(want))
        for i, (d, sm) in enumerate(wmaps):
            assert np.array_equal(two.get_map(i, B.MAP_DOUBLE, stream=stream).view(np.uint32),
                                  d.view(np.uint32))
            assert np.array_equal(two.get_map(i, B.MAP_SINGLE, stream=stream).view(np.uint32),
                                  sm.view(np.uint32))


@pytest.mark.parametrize("flags_name", ["two_streams", "single_stream"])
def test_profile_counts_every_launch(B, flags_name):
    """With BIMP_FLAG_PROFILE every launch of a frame is timed and counted, whether it ran on
    the caller's stream or the context's second one: once per frame for the two NMS passes,
    per pyramid level for the pyramid, and for the cells stage once per inverse sub-batch (the
    scales whose maps become ready together share a launch), each with a finite positive
    duration; no separate clearing of the max slots (each keypoint pass clears them)."""
    import math

    import oracle_lib as O
    w, h, frames = 640, 480, 3
    img = O.synth_frame(w, h, 60, 9)
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    extra = B.FLAG_SINGLE_STREAM if flags_name == "single_stream" else 0
    cs.set_flags(B.FLAG_PROFILE | extra)
    cs.reset_profile()
    want, wn = None, None
    for _ in range(frames):
        got, n = cs.detect(img, cap=1 << 16)
        if want is None:
            want, wn = got, n
        assert n == wn
    prof = cs.profile()
    levels = max(si.level for si in cs.ks)
    count = {k: v["launches"] for k, v in prof.items()}
    # the max slots are cleared by each frame's keypoint pass: back-to-back detections need no
    # separate reset
    assert count.get("k_init", 0) == 0
    assert count["k_pyr_down"] == frames * levels
    inv_rows = sum(v for k, v in count.items() if k.startswith("k_inv_rows"))
    assert count["k_cells"] == inv_rows and frames <= inv_rows <= frames * len(cs.ks)
    for k in ("k_nms_count", "k_nms_write"):
        assert count[k] == frames, k
    assert any(k.startswith("k_inv_cols") for k in prof) and any(k.startswith("k_inv_rows") for k in prof)
    for k, v in prof.items():
        assert math.isfinite(v["total_ms"]) and v["total_ms"] > 0, k


def test_max_slots_across_entry_points(B):
    """The max slots are cleared by each keypoint pass and reset before cells only when left
    holding maxima: detect, filter + cells + keypoints, keypoints again on the same maps, and
    a graph frame in between all give the same keypoints as a fresh context."""
    import oracle_lib as O
    from kpmatch import bits
    w, h = 320, 240
    img = O.synth_frame(w, h, 25, 13)
    ref, rn = B.CudaStuff(w, h, O.BASE_BANK).detect(img)
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    for _ in range(2):
        got, n = cs.detect(img)
        assert n == rn and np.array_equal(bits(got), bits(ref))
    cs.filter(img)
    cs.cells()
    for _ in range(2):  # the second pass recomputes the maxima the first one cleared
        got, n = cs.keypoints()
        assert n == rn and np.array_equal(bits(got), bits(ref))
    cs.filter(img)
    cs.cells()
    cs.set_flags(B.FLAG_GRAPH)
    got, n = cs.detect(img)
    assert n == rn and np.array_equal(bits(got), bits(ref))
    cs.set_flags(0)
    got, n = cs.detect(img)
    assert n == rn and np.array_equal(bits(got), bits(ref))
