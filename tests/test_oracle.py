"""Pins of the CPU oracle (CPU only).

The reference ships no tests, fixtures or golden vectors, so each oracle function is pinned by
what the mathematics fixes: known-answer tables from the reference's own formulas, closed forms,
invariants, an independent FFT (numpy's pocketfft), brute-force circular convolution, and
hand-derived impulse responses chosen so that a dropped term, a wrong sign or offset, or the
wrong rounding mode fails.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- snapTransformSize / lambdas
@pytest.mark.parametrize("n,want", [
    (1080, 1536), (1920, 2048), (540, 1024), (960, 1024), (480, 512), (640, 1024), (240, 256),
    (720, 1024), (1280, 1536), (2160, 2560), (3840, 4096), (512, 512), (320, 512), (100, 128),
    (17, 32), (16, 16), (1025, 1536), (1024, 1024), (1537, 2048), (48, 64), (64, 64),
])
def test_snap_known_answers(n, want):
    # align to 16; a power of two stays; else next power of two if <= 1024; else align to 512
    assert O.snap(n) == want


def test_make_lambdas_log_float_accumulation():
    lam = O.make_lambdas_log(4, 64, 2)
    assert len(lam) == 9
    m = np.float32(2.0 ** 0.5)
    acc = [np.float32(4.0)]
    for _ in range(8):
        acc.append(np.float32(acc[-1] * m))
    assert np.array_equal(lam, np.array(acc, np.float32))
    # the hazard: the third wavelength is 7.9999995, not 8 (float accumulation)
    assert lam[2] == np.float32(7.9999995) and lam[2] < 8


@pytest.mark.parametrize("lam2,p,kW,kX,off,H,W,FH,FW", [
    (4.0, 1, 29, 14, 4, 1080, 1920, 1536, 2048),
    (4.7568283, 1, 33, 16, 4, 1080, 1920, 1536, 2048),
    (5.656854, 1, 39, 19, 5, 1080, 1920, 1536, 2048),
    (6.7271714, 1, 47, 23, 6, 1080, 1920, 1536, 2048),
    (8.0, 2, 29, 14, 4, 540, 960, 1024, 1024),
    (9.513657, 2, 33, 16, 4, 540, 960, 1024, 1024),
    (11.313708, 2, 39, 19, 5, 540, 960, 1024, 1024),
    (13.454343, 2, 47, 23, 6, 540, 960, 1024, 1024),
    (7.9999995, 2, 27, 13, 3, 540, 960, 1024, 1024),  # the makeLambdasLog hazard
    (16.0, 4, 29, 14, 4, 270, 480, 512, 512),
])
def test_scale_parameters(lam2, p, kW, kX, off, H, W, FH, FW):
    s = O.scale_params(np.float32(lam2), 1920, 1080)
    assert (s.pyrstep, s.kW, s.kX, s.off, s.dataH, s.dataW, s.fftH, s.fftW) == \
        (p, kW, kX, off, H, W, FH, FW)
    assert s.kW % 2 == 1 and s.kX == (s.kW - 1) // 2


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 15, 16, 25, 30, 48, 64, 96,
                               120, 128, 1536, 2048, 2560, 3072, 3584])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = O.fft_c2c(x, sign)
    want = np.fft.fft(x) if sign < 0 else np.fft.ifft(x) * n
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


# ---------------------------------------------------------------- Gabor bank
def _bank_scale(lam2=4.7568283):
    return O.scale_params(np.float32(lam2), 320, 240)


@pytest.mark.parametrize("o", range(8))
def test_gabor_symmetry_and_dc(o):
    s = _bank_scale()  # fs = 33.298 but kW = 33
    ke, ko = O.gabor(s, o)
    # even kernel symmetric, odd kernel antisymmetric under (x, y) -> (-x, -y)
    assert np.array_equal(ke, ke[::-1, ::-1])
    assert np.abs(ko + ko[::-1, ::-1]).max() <= 1e-6 * np.abs(ko).max()
    # the raw centre tap is exactly 1, so dc = 1 - centre; dividing the even sum by fs^2
    # (not kW^2) leaves sum(Ke) = dc * (fs^2 - kW^2) instead of 0
    dc = 1.0 - float(ke[s.kX, s.kX])
    resid = float(ke.astype(np.float64).sum())
    want = dc * (float(s.fs) ** 2 - s.kW ** 2)
    assert abs(want) > 1e-4
    assert abs(resid - want) < 0.05 * abs(want)


def test_gabor_axis_closed_forms():
    """theta = 0 oscillates along x (columns); along y the envelope uses gamma, not gamma^2;
    theta = pi/4 oscillates along the row+column diagonal."""
    s = _bank_scale()
    lam, sig = float(s.lam), float(np.float32(s.lam * 0.56))
    ke0, ko0 = O.gabor(s, 0)
    dc = 1.0 - float(ke0[s.kX, s.kX])
    for d in range(1, 6):
        env = np.exp(-d * d / (2 * sig * sig))
        assert abs(ke0[s.kX, s.kX + d] - (env * np.cos(2 * np.pi * d / lam) - dc)) < 2e-6
        assert abs(ko0[s.kX, s.kX + d] - env * np.sin(2 * np.pi * d / lam)) < 2e-6
        assert abs(ke0[s.kX + d, s.kX] - (np.exp(-0.5 * d * d / (2 * sig * sig)) - dc)) < 2e-6
    ke2, ko2 = O.gabor(s, 2)
    dc2 = 1.0 - float(ke2[s.kX, s.kX])
    for d in range(1, 4):
        want = np.exp(-2 * d * d / (2 * sig * sig)) * np.cos(2 * np.pi * np.sqrt(2) * d / lam) - dc2
        assert abs(ke2[s.kX + d, s.kX + d] - want) < 2e-6


@pytest.mark.parametrize("o", [0, 2, 3])
def test_quadrature_spectrum_real_and_one_sided(o):
    s = O.scale_params(np.float32(5.656854), 200, 120)
    ke, ko = O.gabor(s, o)
    a = np.zeros((s.fftH, s.fftW), complex)
    for r in range(s.kW):
        for c in range(s.kW):
            a[(r - s.kX) % s.fftH, (c - s.kX) % s.fftW] += ke[r, c] + 1j * ko[r, c]
    q = np.fft.fft2(a)
    assert np.abs(q.imag).max() < 1e-6 * np.abs(q.real).max()
    ky, kx = np.unravel_index(np.argmax(q.real), q.shape)
    th = o * np.pi / 8
    fy = ((ky + s.fftH // 2) % s.fftH - s.fftH // 2) / s.fftH
    fx = ((kx + s.fftW // 2) % s.fftW - s.fftW // 2) / s.fftW
    # peak at +f0 along (cos theta, sin theta) in (x, y): the odd part is +sin
    assert abs(fx - np.cos(th) / s.lam) < 1.5 / s.fftW + 0.02
    assert abs(fy - np.sin(th) / s.lam) < 1.5 / s.fftH + 0.02


# ---------------------------------------------------------------- pyramid
def test_pyr_down_known_answers():
    img = np.zeros((9, 9), np.uint8)
    img[4, 4] = 128
    d = O.pyr_down(img)
    assert d.shape == (5, 5)
    assert d[2, 2] == 18   # 128 * 36 / 256
    assert d[2, 1] == 3    # 128 * 6 / 256
    assert d[1, 1] == 0    # 128 / 256 = 0.5 rounds half to even -> 0
    assert d[2, 0] == 0
    img = np.zeros((9, 9), np.uint8)
    img[4, 4] = 64
    assert O.pyr_down(img)[1, 2] == 2   # 64 * 1 * 6 / 256 = 1.5 -> 2
    img[4, 5] = 64
    assert O.pyr_down(img)[1, 2] == 2   # (64 * 6 + 64 * 4) / 256 = 2.5 -> 2 (half to even)
    img = np.zeros((9, 9), np.uint8)
    img[2, 4] = 64          # reflect-101 reaches row 2 twice from dst row 0 (taps -2 and +2)
    assert O.pyr_down(img)[0, 2] == 3   # 64 * 2 * 6 / 256 = 3 (reflect would give 1.5 -> 2)
    ramp = np.tile((16 * np.arange(6)).astype(np.uint8), (6, 1))
    assert O.pyr_down(ramp)[0, 0] == 12  # reflect-101 border (reflect gives 7, replicate 6)
    const = np.full((7, 10), 77, np.uint8)
    assert np.all(O.pyr_down(const) == 77)


def test_pyr_down_host_rounding_known_answers():
    """The resize_cpu pyramid (OpenCV host pyrDown, (S + 128) >> 8): half-way sums round up."""
    img = np.zeros((9, 9), np.uint8)
    img[4, 4] = 128
    d = O.pyr_down(img, half_up=True)
    assert d[2, 2] == 18 and d[2, 1] == 3
    assert d[1, 1] == 1    # 128 / 256 = 0.5 -> 1 (half to even gives 0)
    img = np.zeros((9, 9), np.uint8)
    img[4, 4] = 64
    img[4, 5] = 64
    assert O.pyr_down(img, half_up=True)[1, 2] == 3   # 2.5 -> 3 (half to even gives 2)
    img[4, 5] = 0
    assert O.pyr_down(img, half_up=True)[1, 2] == 2   # 1.5 -> 2 either way
    img = np.zeros((9, 9), np.uint8)
    img[2, 4] = 64
    assert O.pyr_down(img, half_up=True)[0, 2] == 3   # reflect-101, as the device pyramid
    const = np.full((7, 10), 255, np.uint8)
    assert np.all(O.pyr_down(const, half_up=True) == 255)   # (65280 + 128) >> 8, no overflow


def _pyr_numpy(img, half_up):
    """Separable [1 4 6 4 1] passes on a reflect-101 padded copy (numpy's 'reflect' mode),
    rounded from the integer sum: an independent statement of both pyramid roundings."""
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = np.pad(img.astype(np.int64), 2, mode="reflect")
    h, w = img.shape
    rows = sum(k[i] * p[i:i + h, :] for i in range(5))
    S = sum(k[j] * rows[:, j:j + w] for j in range(5))[::2, ::2]
    if half_up:
        return ((S + 128) // 256).astype(np.uint8)
    q, r = np.divmod(S, 256)
    return (q + ((r > 128) | ((r == 128) & (q % 2 == 1)))).astype(np.uint8)


@pytest.mark.parametrize("w,h,seed", [(2, 2, 0), (3, 5, 1), (17, 9, 2), (64, 48, 3),
                                      (101, 77, 4)])
def test_pyr_down_both_roundings_match_numpy(w, h, seed):
    img = np.random.default_rng(seed).integers(0, 256, (h, w), dtype=np.uint8)
    for half_up in (False, True):
        assert np.array_equal(O.pyr_down(img, half_up), _pyr_numpy(img, half_up)), half_up
    # the two differ exactly at half-way sums whose floor is even
    a, b = O.pyr_down(img), O.pyr_down(img, half_up=True)
    d = b.astype(int) - a.astype(int)
    assert set(np.unique(d)) <= {0, 1}


def test_oracle_context_resize_cpu_pyramid():
    """bo_set_resize_cpu: the context's levels are the host-rounded chain."""
    w, h = 161, 121
    lam = O.BASE_BANK
    img = O.synth_frame(w, h, 12, 5)
    orc = O.Oracle(w, h, lam)
    orc.set_resize_cpu(True)
    orc.set_image(img)
    for i, s in enumerate(orc.scales):
        want = img
        for _ in range(s.level):
            want = O.pyr_down(want, half_up=True)
        assert np.array_equal(orc.level_image(i), want[:s.dataH, :s.dataW]), i


def test_pad_clamp_known_answers():
    s = O.scale_params(np.float32(4.0), 20, 10)  # kX = 14, FFT 16 x 32
    img = np.arange(200, dtype=np.uint8).reshape(10, 20)
    pd = O.pad_clamp(img, s)
    sc = np.float32(255.0 / 65536.0)
    assert pd.shape == (16, 32)
    assert pd[3, 5] == np.float32(img[3, 5]) * sc
    assert pd[12, 5] == np.float32(img[9, 5]) * sc     # rows H..H+kY-1 repeat row H-1
    assert pd[3, 25] == np.float32(img[3, 19]) * sc    # columns W..W+kX-1 repeat column W-1
    s2 = O.scale_params(np.float32(4.0), 40, 20)       # FFT 32 x 64: rows >= H+kY exist
    img2 = np.arange(800, dtype=np.int64).reshape(20, 40).astype(np.uint8)
    pd2 = O.pad_clamp(img2, s2)
    assert pd2[31, 3] == np.float32(img2[19, 3]) * sc  # 31 < 20 + 14
    assert pd2[10, 60] == np.float32(img2[10, 0]) * sc  # 60 >= 40 + 14 -> column 0


# ---------------------------------------------------------------- filter stage
@pytest.mark.parametrize("lam2,w,h", [(4.0, 40, 24), (5.656854, 40, 24), (8.0, 60, 44)])
def test_fft_filter_equals_brute_force_circular_convolution(lam2, w, h):
    img = O.synth_frame(w, h, 4, 99)
    orc = O.Oracle(w, h, [lam2])
    orc.set_image(img)
    s = orc.scales[0]
    lev = orc.level_image(0)
    maps = orc.filter(0)
    for o in (0, 3, 6):
        e, od = O.brute_filter(lev, s, o)
        c = np.sqrt(e * e + od * od)
        assert np.abs(maps[o] - c).max() <= 1e-6 * c.max()
        # the same convolution by numpy's FFT, kernel padded with wrap-around
        ke, ko = O.gabor(s, o)
        pad = np.zeros((s.fftH, s.fftW), complex)
        for r in range(s.kW):
            for cc in range(s.kW):
                pad[(r - s.kX) % s.fftH, (cc - s.kX) % s.fftW] += ke[r, cc] + 1j * ko[r, cc]
        z = np.fft.ifft2(np.fft.fft2(O.pad_clamp(lev, s).astype(np.float64)) * np.fft.fft2(pad))
        assert np.abs(np.abs(z) - c).max() <= 1e-9 * c.max()


# ---------------------------------------------------------------- cells (impulse responses)
def _cells_ctx(w=40, h=36, lam2=4.0):
    orc = O.Oracle(w, h, [lam2])
    s = orc.scales[0]
    return orc, s, np.zeros((8, s.fftH, s.fftW), np.float32)


def test_offsets_half_away_and_swap():
    assert [O.end_stop_offsets(4.0, o) for o in range(8)] == \
        [(2, 0), (2, 1), (2, 2), (1, 2), (0, 2), (-1, 2), (-2, 2), (-2, 1)]
    assert [O.end_stop_offsets(np.float32(6.7271714), o) for o in range(8)] == \
        [(4, 0), (4, 2), (3, 3), (2, 4), (0, 4), (-2, 4), (-3, 3), (-4, 2)]
    for o in range(8):
        a, b = O.end_stop_offsets(4.0, o)
        assert O.inhibition_offsets(4.0, o) == (b, a)


def test_cells_impulse_orientation0():
    orc, s, maps = _cells_ctx()
    y0, x0 = 16, 18
    maps[0, y0, x0] = 1.0
    D, S = orc.cells(0, maps)
    # o = 0, lambda = 4: end-stop (a, b) = (2, 0); inhibition (a', b') = (0, 2), half (0, 1)
    # single-stopped at (y0+2, x0): C(y-2, x) - C(y+2, x) = 1
    assert S[y0 + 2, x0] == 1.0
    assert S[y0 - 2, x0] == 0.0
    # double-stopped at the impulse: 1; radial at the impulse (partner map 4 empty): 1 + 1 = 2
    assert D[y0, x0] == np.float32(1.0 - 2.0)
    # tangential reads the minus side only: pixel (y0, x0+2) sees C(y0, x0) - 0 = 1 ...
    assert D[y0, x0 + 2] == -1.0 and S[y0, x0 + 2] == -1.0
    # ... and pixel (y0, x0-2), whose plus side is the impulse, sees nothing
    assert D[y0, x0 - 2] == 0.0 and S[y0, x0 - 2] == 0.0
    # double-stopped neighbours at (y0 +- 4, x0): C - 0.5 * 1 < 0 -> 0; no inhibition there
    assert D[y0 + 4, x0] == 0.0


def test_cells_radial_uses_16x_partner_and_half_away_offsets():
    orc, s, maps = _cells_ctx()
    # o = 1 at lambda = 4: (a', b') = (1, 2); half offsets round(0.5) = 1 (half away), 1
    y0, x0 = 16, 18
    maps[1, y0, x0] = 20.0        # centre of orientation 1
    maps[5, y0 + 1, x0 + 1] = 1.0  # partner map o+4 at (+a'', +b'')
    D, S = orc.cells(0, maps)
    # radial: max(0, max(0, 20 - 16*1) + 20 - 16*0) = 24; tangential of o=1 at the centre: 0
    # end-stop of o=1 at the centre: S term 0, D term 20 - 0 - 0 = 20
    # orientation 5 at (y0+1, x0+1): its partner is map 1 at (+-1, +-2)... none -> radial 2*1
    assert D[y0, x0] == np.float32(20.0 - 24.0)
    assert S[y0, x0] == np.float32(0.0 - 24.0)


def test_cells_inclusive_bounds():
    orc, s, maps = _cells_ctx()
    H, W = s.dataH, s.dataW
    maps[:] = 1.0
    maps[0, :, :] = 2.0
    D, S = orc.cells(0, maps)
    # constant maps: end-stop S terms are 0, D terms C - C = 0, tangential 0, radial
    # max(0, C_o - 16 C_o2) = 0 for every orientation: all zero inside
    assert np.all(D == 0) and np.all(S == 0)
    # o = 0 end-stop validity: y - 4 >= 0 and y + 4 <= H, inclusive of row H.  The
    # single-stopped term at (y, x) is C(y-2, x) - C(y+2, x); with an impulse at (y-2, x)
    # and nothing else in reach, S(y, x) is 1 exactly when orientation 0 is valid there.
    for y, valid in [(H - 4, True), (H - 3, False), (4, True), (3, False)]:
        maps[:] = 0
        maps[0, y - 2, 10] = 1.0
        D, S = orc.cells(0, maps)
        assert S[y, 10] == (1.0 if valid else 0.0), (y, valid)


# ---------------------------------------------------------------- NMS
def _nms_ctx(w=40, h=36):
    return O.Oracle(w, h, [4.0])


def test_nms_parabola_and_fields():
    orc = _nms_ctx()
    D = np.zeros((36, 40), np.float32)
    y, x = 10, 12
    D[y, x] = 5
    D[y, x - 1], D[y, x + 1], D[y - 1, x], D[y + 1, x] = 3, 4, 2, 1
    kps, n = orc.nms(0, D)
    assert n == 1
    k = kps[0]
    assert k["x"] == np.float32(np.float32(12) + np.float32(np.float32(-0.5) / np.float32(-3.0)))
    assert k["y"] == np.float32(np.float32(10) + np.float32(np.float32(0.5) / np.float32(-7.0)))
    assert (k["size"], k["angle"], k["response"], k["octave"], k["class_id"]) == (4, -1, 5, 4, 1)


def test_nms_strict_max_threshold_and_border():
    orc = _nms_ctx()
    D = np.zeros((36, 40), np.float32)
    D[10, 10] = 10            # keypoint
    D[20, 20] = 10; D[20, 21] = 10   # plateau: neither is a strict maximum
    D[15, 30] = 1.0           # 1.0 < 0.104 * 10 = 1.04: below threshold
    D[15, 5] = 8              # x = 5 > off = 4: inside
    D[25, 4] = 8              # x = off: excluded
    D[31, 15] = 8             # y = H - off - 1 = 31: inside
    D[32, 25] = 8             # y = H - off: excluded
    kps, n = orc.nms(0, D)
    got = sorted((int(round(k["y"])), int(round(k["x"]))) for k in kps)
    assert got == [(10, 10), (15, 5), (31, 15)]
    # row-major order within the scale
    assert [(int(round(k["y"])), int(round(k["x"]))) for k in kps] == [(10, 10), (15, 5), (31, 15)]


def test_nms_threshold_uses_double_product():
    orc = _nms_ctx()
    D = np.zeros((36, 40), np.float32)
    D[10, 10] = 100.0
    t = np.float32(100.0 * 0.104)
    D[20, 20] = np.nextafter(t, np.float32(np.inf))   # just above: kept
    D[25, 25] = t                                      # equal: not strictly greater
    kps, n = orc.nms(0, D)
    assert n == 2


# ---------------------------------------------------------------- synthetic frames
def test_synthetic_frame_hashes():
    with open(os.path.join(GOLDEN, "synth_hashes.json")) as f:
        want = json.load(f)
    for key, h in want.items():
        w, hh, n, seed = (int(v) for v in key.split("_"))
        img = O.synth_frame(w, hh, n, seed)
        assert hashlib.sha256(img.tobytes()).hexdigest() == h, key


def test_cfg0_oracle_keypoints_fixture():
    """Regression fixture of the oracle itself on the 640x480 workload (FP64 filter path)."""
    want = np.load(os.path.join(GOLDEN, "cfg0_keypoints.npy"), allow_pickle=False)
    img = O.synth_frame(640, 480, 40, 0)
    orc = O.Oracle(640, 480, O.BASE_BANK)
    kps, n = orc.detect(img)
    assert n == len(want)
    got = np.ascontiguousarray(kps).view(np.float32).reshape(n, 7)
    assert np.abs(got - want).max() <= 1e-3


# ---------------------------------------------------------------- CPU keypoint mode (Q18)
def _cpu_mode_maps(w=40, h=36):
    orc = O.Oracle(w, h, [4.0])
    s = orc.scales[0]
    return orc, s, np.zeros((h, w), np.float32), np.zeros((h, w), np.float32)


def test_cpu_mode_inclusive_border_and_single_keypoint():
    orc, s, D, S = _cpu_mode_maps()
    off = s.off  # 4
    # samples of 10 - (t - 0.25)^2 in x and 10 - (t + 0.5)^2 in y at t = -1, 0, 1 (exact in
    # float): the parabola fit puts the peak at +0.25 and -0.5
    D[off, off] = 9.9375        # on the inclusive border: kept (the GPU NMS needs off < x)
    D[off, off - 1] = 8.4375
    D[off, off + 1] = 9.4375
    D[off - 1, off] = 9.75 - 0.0  # 10 - 0.25
    D[off + 1, off] = 7.75        # 10 - 2.25
    D[off, off] = 9.9375
    kps, n = orc.nms_cpu_mode(0, D, S)
    assert n == 1
    k = kps[0]
    assert k["x"] == off + 0.25
    # y samples: 9.75, 9.9375 (centre), 7.75 are not one parabola with the centre of x; the fit
    # of (9.75, 9.9375, 7.75): b = -1, a = -1.1875 -> peak -0.42105...
    assert abs(k["y"] - (off - 1.0 / 2.375)) < 1e-6
    assert k["class_id"] == -1 and k["octave"] == 4 and k["size"] == 4.0
    assert k["response"] == np.float32(9.9375)
    gkps, gn = orc.nms(0, D)
    assert gn == 0  # the GPU rule excludes x == off


def test_cpu_mode_adds_single_stopped_offsets_and_uses_the_double_threshold():
    orc, s, D, S = _cpu_mode_maps()
    y, x = 15, 20
    D[y, x], D[y, x - 1], D[y, x + 1] = 10.0, 2.0, 6.0
    S[y, x], S[y, x - 1], S[y, x + 1] = 2.0, 1.0, 0.5   # S peak above 0.104 * max(D) only
    S[y - 1, x], S[y + 1, x] = 0.25, 0.75
    kps, n = orc.nms_cpu_mode(0, D, S)
    assert n == 1
    kd, _ = orc.nms_cpu_mode(0, D, np.zeros_like(S))
    # with the S peak the position moves by the S-map parabola offsets as well
    assert kps[0]["x"] != kd[0]["x"] and kps[0]["y"] != kd[0]["y"]
    # an S peak below 0.104 * max(D) does not move it
    S2 = S.copy()
    S2[y, x] = 1.03
    kps2, _ = orc.nms_cpu_mode(0, D, S2)
    assert kps2[0]["x"] == kd[0]["x"] and kps2[0]["y"] == kd[0]["y"]


def test_cpu_mode_plateau_and_outside_neighbours():
    orc, s, D, S = _cpu_mode_maps()
    D[10, 10] = D[10, 11] = 5.0   # plateau: neither is a strict maximum
    D[20, 20] = 5.0
    kps, n = orc.nms_cpu_mode(0, D, S)
    assert n == 1 and kps[0]["response"] == 5.0
