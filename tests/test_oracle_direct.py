"""Pins of the oracle's precision form (bo_direct_taps, bo_direct_filter_scale).

The precision form convolves with taps rounded once to BF16 or FP16.  Its oracle is pinned by
independent routes: numpy's own float16 conversion and a bit-level BF16 rounding for the taps,
a numpy FFT circular convolution with the same rounded taps for the indexing (a flipped
kernel, a wrong centre or a lost box term fails it), and the FP64 brute-force filter with
unrounded taps for the size of the rounding error.
"""
import numpy as np
import pytest

import oracle_lib as O

ISCALE = 255.0 / 65536.0


def bf16_round(x32):
    """Round float32 values to BF16 (nearest-even on the upper 16 bits), as float64."""
    u = np.asarray(x32, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _raw_bank(s):
    """Unrounded filters in filter order (2o even, 2o+1 odd), FP64."""
    return np.stack([k for o in range(8) for k in O.gabor(s, o)]).astype(np.float64)


@pytest.mark.parametrize("fp16", [False, True])
@pytest.mark.parametrize("lam2", [4.0, 5.656854])
def test_full_taps_are_the_rounded_gabor_bank(lam2, fp16):
    s = O.scale_params(lam2, 64, 48)
    ext, fscale, dc, taps = O.direct_taps(s, 0.0, fp16)
    assert ext == s.kW
    raw = _raw_bank(s)
    mx = np.abs(raw).max()
    assert fscale == float(np.float32(mx))
    q = (raw / mx).astype(np.float32)
    want = q.astype(np.float16).astype(np.float64) if fp16 else bf16_round(q)
    assert np.array_equal(taps, want)
    # BF16 keeps 8 significant bits, FP16 11 (and flushes below 2^-24)
    rel = np.abs(taps - raw / mx)
    assert rel.max() <= (2.0 ** -9 if not fp16 else 2.0 ** -12)
    # the box coefficient only absorbs the rounding's sum: tiny against the taps
    assert np.abs(dc).max() * s.kW ** 2 <= s.kW * (2.0 ** -8 if not fp16 else 2.0 ** -11) * fscale


def test_truncated_taps_are_the_carrier():
    s = O.scale_params(5.656854, 64, 48)
    ext, fscale, dc, taps = O.direct_taps(s, 6.0, False)
    sigma = float(np.float32(np.float64(np.float32(5.656854)) * 0.56))
    want_ext = int(np.ceil(6.0 * sigma))
    want_ext += 1 - want_ext % 2
    assert ext == want_ext == 21 and ext < s.kW
    R, kX = ext // 2, s.kX
    inside = np.zeros((s.kW, s.kW), bool)
    inside[kX - R:kX + R + 1, kX - R:kX + R + 1] = True
    assert not taps[:, ~inside].any()
    for o in range(8):
        ke, ko = O.gabor(s, o)
        # Ke sums to dcK * (fs^2 - kW^2), dcK the offset bo_gabor subtracted
        dck = ke.astype(np.float64).sum() / (float(s.fs) ** 2 - s.kW ** 2)
        carrier = (ke.astype(np.float64) + dck)[inside]
        assert np.abs(taps[2 * o][inside] * fscale - carrier).max() <= 2.0 ** -8 * fscale
        assert np.abs(taps[2 * o + 1][inside] * fscale - ko[inside]).max() <= 2.0 ** -9 * fscale


@pytest.mark.parametrize("nsig,fp16", [(0.0, False), (0.0, True), (6.0, False), (8.0, True)])
def test_constant_image_gets_the_full_kernel_response(nsig, fp16):
    """The box coefficient removes the DC leak of rounding and truncation exactly."""
    w, h, m = 64, 48, 137
    orc = O.Oracle(w, h, [5.656854])
    orc.set_image(np.full((h, w), m, np.uint8))
    s = orc.scales[0]
    maps = orc.direct_filter(0, nsig, fp16)
    ref = orc.filter(0)
    for o in range(8):
        ke, ko = O.gabor(s, o)
        want = ISCALE * m * np.hypot(ke.astype(np.float64).sum(), ko.astype(np.float64).sum())
        assert np.allclose(maps[o], want, rtol=1e-6, atol=1e-12)
        assert np.allclose(ref[o], want, rtol=1e-3, atol=1e-7)


def _fft_route(lev, s, taps, fscale, dc):
    """Same maps by numpy FFT circular convolution of the uint8 padded image."""
    P = O.pad_clamp(lev, s).astype(np.float64) / np.float64(np.float32(ISCALE))
    P = np.rint(P)  # back to the uint8 samples
    FP = np.fft.fft2(P)

    def conv(k):
        pad = np.zeros((s.fftH, s.fftW))
        for r in range(s.kW):
            for c in range(s.kW):
                pad[(r - s.kX) % s.fftH, (c - s.kX) % s.fftW] += k[r, c]
        return np.real(np.fft.ifft2(FP * np.fft.fft2(pad)))

    box = conv(np.ones((s.kW, s.kW)))
    out = []
    for o in range(8):
        e = fscale * conv(taps[2 * o]) - dc[2 * o] * box
        od = fscale * conv(taps[2 * o + 1]) - dc[2 * o + 1] * box
        out.append(ISCALE * np.sqrt(e * e + od * od))
    return np.stack(out)


@pytest.mark.parametrize("lam2,nsig,fp16", [(4.0, 0.0, False), (5.656854, 0.0, True),
                                            (5.656854, 6.0, False), (4.0, 8.0, True)])
def test_direct_filter_equals_fft_route_with_rounded_taps(lam2, nsig, fp16):
    w, h = 64, 48
    img = O.synth_frame(w, h, 5, 1234)
    orc = O.Oracle(w, h, [lam2])
    orc.set_image(img)
    s = orc.scales[0]
    lev = orc.level_image(0)
    ext, fscale, dc, taps = O.direct_taps(s, nsig, fp16)
    maps = orc.direct_filter(0, nsig, fp16)
    want = _fft_route(lev, s, taps, fscale, dc)
    assert np.abs(maps - want).max() <= 1e-6 * want.max()


@pytest.mark.parametrize("lam2", [4.0, 5.656854])
@pytest.mark.parametrize("nsig,fp16,tol", [(0.0, False, 3e-3), (0.0, True, 1e-3),
                                           (8.0, False, 6e-3), (8.0, True, 4e-3)])
def test_direct_filter_error_against_fp64_filter(lam2, nsig, fp16, tol):
    w, h = 64, 48
    img = O.synth_frame(w, h, 5, 77)
    orc = O.Oracle(w, h, [lam2])
    orc.set_image(img)
    s = orc.scales[0]
    lev = orc.level_image(0)
    maps = orc.direct_filter(0, nsig, fp16)
    for o in range(8):
        e, od = O.brute_filter(lev, s, o)
        c = np.sqrt(e * e + od * od)
        assert np.abs(maps[o] - c).max() <= tol * c.max()
