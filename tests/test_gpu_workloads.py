"""GPU parity at the workloads' own sizes and on the edge cases of the frame geometry.

- Precision form on the 1920x1080 frame (300 shapes, seed 1) in BF16 and FP16, for the
  shipped form (every scale direct, 8 sigma) and for the split lambda <= 8 direct at 8 sigma,
  the rest FFT-FP32: per-scale complex-cell map error against the FP64 oracle within 1e-2
  (direct scales) or 1e-4 (FFT scales), keypoint agreement >= 99%.
- The reference Context's own bank, makeLambdasLog(4, 64, 2) (bimp.cpp:8): 9 scales, pyramid
  level 4, the float-accumulated 7.9999995 wavelength (kW 27, off 3), at 640x480 and 1920x1080.
- Odd frame sizes (641x479, 1919x1079): pyramid levels of (n+1)/2 against the floor(n/2) data
  crop of the upper scales (cuda_bimp.cu:473-486).
- Every pyramid level bit-exact against the oracle's pyrDown (5-tap binomial, reflect-101,
  round half to even), on even and odd sizes.
- The resize_cpu pyramid (BIMP_FLAG_RESIZE_CPU: host pyrDown rounding, half up): every level
  bit-exact against the oracle's host-rounded chain, and detection end to end against the
  oracle built on that pyramid.
- The differentiation form at its workload: 16 FP32 512x512 images, 8 scales.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from kpmatch import agreement, rel_err

pytestmark = pytest.mark.gpu

MAP_TOL = 1e-4
PREC_MAP_TOL = 1e-2


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    O.lib().bo_set_threads(min(16, os.cpu_count() or 1))
    return bimp_amd


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ precision form at 1080p
@pytest.fixture(scope="module")
def frame1080(B):
    """The 1080p frame, the FP64 oracle's complex-cell maps of every scale and its keypoints."""
    w, h = 1920, 1080
    img = B.synth_frame(w, h, 300, 1)
    orc = O.Oracle(w, h, O.BASE_BANK, fp64=True)
    orc.set_image(img)
    maps, kps = [], []
    for i in range(len(orc.scales)):
        m = orc.filter(i)
        D, _ = orc.cells(i, m)
        k, _ = orc.nms(i, D, cap=1 << 16)
        maps.append(m)
        kps.append(k)
    return img, maps, np.concatenate(kps)


@pytest.mark.parametrize("max_lambda", [64.0, 8.0], ids=["all_direct", "lambda8_split"])
@pytest.mark.parametrize("fmt", ["bf16", "fp16"])
def test_precision_form_1080p_against_fp64(B, frame1080, fmt, max_lambda):
    img, maps, ref_kps = frame1080
    w, h = 1920, 1080
    opt = B.default_options()
    opt.precision = B.PRECISION_FP16 if fmt == "fp16" else B.PRECISION_BF16
    opt.direct_sigmas = 8.0
    opt.direct_max_lambda = max_lambda
    cs = B.CudaStuff(w, h, O.BASE_BANK, opt)
    got, n = cs.detect(img, cap=1 << 16)
    ndirect = 0
    for i, si in enumerate(cs.ks):
        direct = si.lambda2 <= max_lambda
        assert (si.precision != B.PRECISION_FP32) == direct, i
        ndirect += direct
        e = max(rel_err(cs.get_map(i, B.MAP_COMPLEX, o), maps[i][o, :si.map_h, :si.map_w])
                for o in range(8))
        _log(f"{fmt} max_lambda {max_lambda} scale {i} lambda {si.lambda2:.4f} "
             f"{'direct ext %d' % si.tap_extent if direct else 'FFT'} map err {e:.2e}")
        assert e <= (PREC_MAP_TOL if direct else MAP_TOL), (i, e)
    assert ndirect == (8 if max_lambda >= 64 else 5)
    a = agreement(got, ref_kps)
    _log(f"{fmt} max_lambda {max_lambda}: keypoints {n} oracle {len(ref_kps)} agreement {a:.4f}")
    assert a >= 0.99


# ------------------------------------------------------------- default bank and odd sizes
def _end_to_end(B, w, h, lam, nshapes, seed, resize_cpu=False):
    cs = B.CudaStuff(w, h, lam)
    orc = O.Oracle(w, h, lam)
    if resize_cpu:
        cs.set_flags(B.FLAG_RESIZE_CPU)
        orc.set_resize_cpu(True)
    img = B.synth_frame(w, h, nshapes, seed)
    kps, m, Ds, Ss = orc.detect(img, cap=1 << 16, want_maps=True)
    got, n = cs.detect(img, cap=1 << 16)
    for i, si in enumerate(cs.ks):
        s = orc.scales[i]
        assert (si.data_h, si.data_w, si.fft_h, si.fft_w) == (s.dataH, s.dataW, s.fftH, s.fftW)
        assert si.kernel_size == s.kW and si.border == s.off
        eD = rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i])
        eS = rel_err(cs.get_map(i, B.MAP_SINGLE), Ss[i])
        _log(f"{w}x{h} scale {i} lambda {si.lambda2:.7g} level {si.level} "
             f"data {si.data_w}x{si.data_h} D {eD:.1e} S {eS:.1e}")
        assert eD < MAP_TOL and eS < MAP_TOL, i
    a = agreement(got, kps)
    _log(f"{w}x{h}: keypoints {n} oracle {m} agreement {a:.4f}")
    assert n == m or a >= 0.99
    assert a >= 0.99
    return cs


@pytest.mark.parametrize("w,h,nshapes,seed", [(640, 480, 40, 0), (1920, 1080, 300, 1)])
def test_reference_default_bank(B, w, h, nshapes, seed):
    lam = B.make_lambdas_log(4, 64, 2)
    assert len(lam) == 9 and lam[2] == np.float32(7.9999995)
    cs = _end_to_end(B, w, h, lam, nshapes, seed)
    assert cs.ks[2].kernel_size == 27 and cs.ks[2].border == 3
    assert max(si.level for si in cs.ks) == 4


@pytest.mark.parametrize("w,h,nshapes,seed", [(641, 479, 40, 5), (1919, 1079, 300, 6),
                                              (321, 241, 25, 7)])
def test_odd_frame_sizes(B, w, h, nshapes, seed):
    _end_to_end(B, w, h, O.BASE_BANK, nshapes, seed)


def _ties(img):
    """Number of half-way sums S = 128 (mod 256) of the 5x5 binomial on a level (reflect-101)."""
    k = np.array([1, 4, 6, 4, 1], np.int64)
    p = np.pad(img.astype(np.int64), 2, mode="reflect")
    h, w = img.shape
    rows = sum(k[i] * p[i:i + h, :] for i in range(5))
    S = sum(k[j] * rows[:, j:j + w] for j in range(5))
    return int(np.count_nonzero(S[::2, ::2] % 256 == 128))


@pytest.mark.parametrize("w,h", [(640, 480), (641, 479), (1919, 1079), (1920, 1080), (1283, 723)])
def test_pyramid_levels_bit_exact(B, w, h):
    lam = B.make_lambdas_log(4, 64, 2)
    cs = B.CudaStuff(w, h, lam)
    img = B.synth_frame(w, h, 60, 40 + w)
    cs.filter(img)
    want = img
    ties = 0
    for level in range(1, max(si.level for si in cs.ks) + 1):
        ties += _ties(want)
        want = O.pyr_down(want)
        got = cs.get_level(level)
        assert got.shape == want.shape, level
        bad = np.count_nonzero(got != want)
        _log(f"{w}x{h} level {level} {got.shape[1]}x{got.shape[0]}: {bad} differing pixels")
        assert bad == 0, level
    _log(f"{w}x{h}: {ties} half-way sums across the levels")
    assert ties > 0


@pytest.mark.parametrize("w,h,nshapes,seed", [(641, 479, 40, 8), (1919, 1079, 300, 9)])
def test_resize_cpu_pyramid(B, w, h, nshapes, seed):
    """BIMP_FLAG_RESIZE_CPU (the reference's resize_cpu, cuda_bimp.cu:649-687): levels equal the
    host-rounded oracle chain and differ from the device rounding where a half-way sum has an
    even floor; the default bank's detection matches the oracle on that pyramid."""
    lam = B.make_lambdas_log(4, 64, 2)
    cs = _end_to_end(B, w, h, lam, nshapes, seed, resize_cpu=True)
    img = B.synth_frame(w, h, nshapes, seed)
    want, moved = img, 0
    for level in range(1, max(si.level for si in cs.ks) + 1):
        even = O.pyr_down(want)
        want = O.pyr_down(want, half_up=True)
        got = cs.get_level(level)
        bad = np.count_nonzero(got != want)
        moved += int(np.count_nonzero(want != even))
        _log(f"{w}x{h} resize_cpu level {level}: {bad} differing pixels")
        assert bad == 0, level
    _log(f"{w}x{h}: {moved} level pixels where the two roundings differ")
    assert moved > 0


# ------------------------------------------------------- differentiation form, 16 x 512^2
def test_diff_form_workload_16x512(B):
    """Forward of all 16 images against the FP64 restatement (1e-4 per scale, images 0 and
    15); image and parameter gradients of image 0 against the FP64 reference gradient (1e-3);
    central differences (h = 1e-3) of the FP64 reference loss along a smooth image direction
    and along each parameter against the adjoint's directional derivatives (5e-3).  The FP32
    loss cannot carry this check at this size: its rounding (~1e-7 of a ~4e4 loss) divided by
    2h is a few percent of the derivative."""
    import torch
    from bimp_amd.autograd import BimpResponses
    from diff_ref import DiffRef
    size, n = 512, 16
    lam = O.BASE_BANK
    R = BimpResponses(size, size, lam)
    ref = DiffRef(size, size, lam)
    imgs = np.stack([B.synth_frame(size, size, 40, 2000 + i) for i in range(n)]).astype(np.float32) / 255.0
    x = torch.tensor(imgs, device="cuda")
    p = torch.tensor([0.56, 0.5], dtype=torch.float64)
    with torch.no_grad():
        target = R(x, torch.tensor([0.56, 0.6], dtype=torch.float64))
        resp = R(x, p).cpu().numpy()

    def split(v):
        out, o = [], 0
        for s in ref.scales:
            k = s.dataH * s.dataW
            out.append(v[o:o + k].reshape(s.dataH, s.dataW))
            o += k
        return out

    for b in (0, n - 1):
        want = ref.forward(imgs[b].astype(np.float64))
        for i, (g, wv) in enumerate(zip(split(resp[b]), want)):
            e = np.abs(g - wv).max() / np.abs(wv).max()
            _log(f"image {b} scale {i} response rel err {e:.2e}")
            assert e <= 1e-4, (b, i)
    # gradients of the batch loss; image 0's share against the FP64 reference
    xi = x.detach().requires_grad_(True)
    pp = p.clone().requires_grad_(True)
    ((R(xi, pp) - target) ** 2).sum().backward()
    x0 = x[:1].detach().requires_grad_(True)
    p0 = p.clone().requires_grad_(True)
    ((R(x0, p0) - target[:1]) ** 2).sum().backward()
    gx = x0.grad[0].cpu().numpy().astype(np.float64)
    gp = p0.grad.numpy()
    assert np.array_equal(x0.grad[0].cpu().numpy(), xi.grad[0].cpu().numpy())
    img0 = imgs[0].astype(np.float64)
    tg = split(target[0].cpu().numpy().astype(np.float64))
    rl, rg, rth = ref.gradient(img0, tg)
    e_img = np.linalg.norm(gx - rg) / np.linalg.norm(rg)
    e_p = np.abs(gp - rth) / np.abs(rth)
    _log(f"image grad rel L2 {e_img:.2e}; param grads {gp} ref {rth} rel {e_p}")
    assert e_img <= 1e-3 and np.all(e_p <= 1e-3)

    def L(im, ratio=0.56, gamma=0.5):
        r = ref.forward(im, ratio=ratio, gamma=gamma)
        return sum(float(((a - t) ** 2).sum()) for a, t in zip(r, tg))

    h = 1e-3
    yy, xx = np.mgrid[0:size, 0:size]
    u = np.sin(2 * np.pi * xx / size) * np.cos(2 * np.pi * yy / size)
    fd = (L(img0 + h * u) - L(img0 - h * u)) / (2 * h)
    an = float((gx * u).sum())
    _log(f"smooth direction: FP64 central difference {fd:.6g} adjoint {an:.6g} "
         f"rel {abs(fd - an) / abs(an):.2e}")
    assert abs(fd - an) <= 5e-3 * abs(an)
    for k, name in enumerate(("ratio", "gamma")):
        a = {"ratio": 0.56, "gamma": 0.5}
        b = dict(a)
        a[name] += h
        b[name] -= h
        fdp = (L(img0, **a) - L(img0, **b)) / (2 * h)
        _log(f"{name}: FP64 central difference {fdp:.6g} adjoint {gp[k]:.6g} "
             f"rel {abs(fdp - gp[k]) / abs(gp[k]):.2e}")
        assert abs(fdp - gp[k]) <= 5e-3 * abs(gp[k])
    R.close()
