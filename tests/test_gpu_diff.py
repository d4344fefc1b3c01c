"""GPU checks of the differentiation form (bimp_diff_* and the libtorch autograd Function).

Bars: responses within 1e-4 of the FP64 restatement (tests/diff_ref.py); image and parameter
gradients within 1e-3 (relative L2 / relative) of its FP64 reference gradient; central
differences of the FP32 loss (h = 1e-3) along a smooth image direction and along each
parameter within 1e-2 of the adjoint's directional derivatives."""
import numpy as np
import pytest

import oracle_lib as O
from diff_ref import DiffRef

pytestmark = pytest.mark.gpu

W, H = 64, 48
LAMS = (4.0, 5.656854, 8.0)


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible to torch")
    return torch


@pytest.fixture(scope="module")
def R(T):
    from bimp_amd.autograd import BimpResponses
    return BimpResponses(W, H, LAMS)


def _img(seed=3):
    return O.synth_frame(W, H, 6, seed).astype(np.float32) / 255.0


def _split(ref, v):
    out, o = [], 0
    for s in ref.scales:
        n = s.dataH * s.dataW
        out.append(v[o:o + n].reshape(s.dataH, s.dataW))
        o += n
    return out


def test_forward_matches_fp64_restatement(T, R):
    img = _img()
    ref = DiffRef(W, H, LAMS)
    want = ref.forward(img.astype(np.float64))
    x = T.tensor(img[None], device="cuda")
    p = T.tensor([0.56, 0.5], dtype=T.float64)
    got = R(x, p).cpu().numpy()[0]
    for i, (g, w) in enumerate(zip(_split(ref, got), want)):
        e = np.abs(g - w).max() / np.abs(w).max()
        print(f"scale {i} response rel err {e:.2e}", flush=True)
        assert e <= 1e-4


def _loss_and_grads(T, R, img, target, params):
    x = T.tensor(img, device="cuda", requires_grad=True)
    p = T.tensor(params, dtype=T.float64, requires_grad=True)
    resp = R(x, p)
    loss = ((resp - target) ** 2).sum()
    loss.backward()
    return float(loss), x.grad.cpu().numpy(), p.grad.numpy()


def test_gradients_match_fp64_reference(T, R):
    img = _img()
    ref = DiffRef(W, H, LAMS)
    targets = ref.forward(img.astype(np.float64), gamma=0.6)
    tgt = T.tensor(np.concatenate([t.ravel() for t in targets])[None].astype(np.float32), device="cuda")
    loss, gx, gp = _loss_and_grads(T, R, img[None], tgt, [0.56, 0.5])
    rl, rg, rth = ref.gradient(img.astype(np.float64), targets)
    e_img = np.linalg.norm(gx[0] - rg) / np.linalg.norm(rg)
    e_p = np.abs(gp - rth) / np.abs(rth)
    print(f"loss {loss:.6g} ref {rl:.6g}; image grad rel L2 {e_img:.2e}; params {gp} ref {rth} "
          f"rel {e_p}", flush=True)
    assert abs(loss - rl) <= 1e-3 * rl
    assert e_img <= 1e-3
    assert np.all(e_p <= 1e-3)


def test_central_differences_h1e3(T, R):
    img = _img(5)
    ref = DiffRef(W, H, LAMS)
    targets = ref.forward(img.astype(np.float64), gamma=0.6)
    tgt = T.tensor(np.concatenate([t.ravel() for t in targets])[None].astype(np.float32), device="cuda")
    _, gx, gp = _loss_and_grads(T, R, img[None], tgt, [0.56, 0.5])
    h = 1e-3

    def L(im, p):
        with T.no_grad():
            r = R(T.tensor(im, device="cuda"), T.tensor(p, dtype=T.float64))
            return float(((r.double() - tgt.double()) ** 2).sum())

    # a smooth image direction: a per-pixel random one at this step flips relu branches of
    # the stencil everywhere and measures the kinks, not the derivative (about 15% in FP64)
    yy, xx = np.mgrid[0:H, 0:W]
    u = (np.sin(2 * np.pi * xx / W) * np.cos(2 * np.pi * yy / H)).astype(np.float32)
    fd = (L((img + h * u)[None], [0.56, 0.5]) - L((img - h * u)[None], [0.56, 0.5])) / (2 * h)
    an = float((gx[0] * u).sum())
    print(f"image direction: fd {fd:.6g} adjoint {an:.6g}", flush=True)
    assert abs(fd - an) <= 1e-2 * abs(an)
    for k in range(2):
        pp, pm = [0.56, 0.5], [0.56, 0.5]
        pp[k] += h
        pm[k] -= h
        fdp = (L(img[None], pp) - L(img[None], pm)) / (2 * h)
        print(f"param {k}: fd {fdp:.6g} adjoint {gp[k]:.6g}", flush=True)
        assert abs(fdp - gp[k]) <= 1e-2 * abs(gp[k])


def test_batch_equals_single_images(T, R):
    imgs = np.stack([_img(s) for s in (3, 4, 5)])
    p = T.tensor([0.56, 0.5], dtype=T.float64)
    tgt = T.zeros((3, R.response_size), device="cuda")
    _, gx, gp = _loss_and_grads(T, R, imgs, tgt, [0.56, 0.5])
    resp = R(T.tensor(imgs, device="cuda"), p).cpu().numpy()
    gsum = np.zeros(2)
    for b in range(3):
        r1 = R(T.tensor(imgs[b:b + 1], device="cuda"), p).cpu().numpy()
        assert np.array_equal(r1[0], resp[b])
        _, g1, gp1 = _loss_and_grads(T, R, imgs[b:b + 1], tgt[b:b + 1], [0.56, 0.5])
        assert np.array_equal(g1[0], gx[b])
        gsum += gp1
    assert np.allclose(gsum, gp, rtol=1e-12, atol=0)


def test_odd_size_upper_level_crop(T):
    """49 x 41 with a level-1 scale: the level image is 25 x 21 and its data the top-left
    24 x 20, so the forward reads and the adjoint writes with the level's pitch."""
    from bimp_amd.autograd import BimpResponses
    w, h = 49, 41
    Ro = BimpResponses(w, h, LAMS)
    img = O.synth_frame(w, h, 6, 11).astype(np.float32) / 255.0
    ref = DiffRef(w, h, LAMS)
    assert ref.scales[-1].level == 1 and ref.scales[-1].dataW == 24
    want = ref.forward(img.astype(np.float64))
    p = T.tensor([0.56, 0.5], dtype=T.float64)
    got = Ro(T.tensor(img[None], device="cuda"), p).cpu().numpy()[0]
    for i, (g, wnt) in enumerate(zip(_split(ref, got), want)):
        e = np.abs(g - wnt).max() / np.abs(wnt).max()
        print(f"scale {i} response rel err {e:.2e}", flush=True)
        assert e <= 1e-4
    targets = ref.forward(img.astype(np.float64), gamma=0.6)
    tgt = T.tensor(np.concatenate([t.ravel() for t in targets])[None].astype(np.float32),
                   device="cuda")
    loss, gx, gp = _loss_and_grads(T, Ro, img[None], tgt, [0.56, 0.5])
    rl, rg, rth = ref.gradient(img.astype(np.float64), targets)
    e_img = np.linalg.norm(gx[0] - rg) / np.linalg.norm(rg)
    e_p = np.abs(gp - rth) / np.abs(rth)
    print(f"image grad rel L2 {e_img:.2e}; params rel {e_p}", flush=True)
    assert e_img <= 1e-3
    assert np.all(e_p <= 1e-3)


def test_4096_wide_transforms(T):
    """3700 x 40 with one scale: 4096-point row transforms (two wavefronts per transform) in the
    forward, the recomputation and the adjoint, against the FP64 restatement."""
    from bimp_amd.autograd import BimpResponses
    w, h, lams = 3700, 40, (4.0,)
    Ro = BimpResponses(w, h, lams)
    ref = DiffRef(w, h, lams)
    assert ref.scales[0].fftW == 4096
    img = O.synth_frame(w, h, 30, 12).astype(np.float32) / 255.0
    want = ref.forward(img.astype(np.float64))
    p = T.tensor([0.56, 0.5], dtype=T.float64)
    got = Ro(T.tensor(img[None], device="cuda"), p).cpu().numpy()[0]
    e = np.abs(got - want[0].ravel()).max() / np.abs(want[0]).max()
    print(f"response rel err {e:.2e}", flush=True)
    assert e <= 1e-4
    targets = ref.forward(img.astype(np.float64), gamma=0.6)
    tgt = T.tensor(np.concatenate([t.ravel() for t in targets])[None].astype(np.float32),
                   device="cuda")
    _, gx, gp = _loss_and_grads(T, Ro, img[None], tgt, [0.56, 0.5])
    _, rg, rth = ref.gradient(img.astype(np.float64), targets)
    e_img = np.linalg.norm(gx[0] - rg) / np.linalg.norm(rg)
    e_p = np.abs(gp - rth) / np.abs(rth)
    print(f"image grad rel L2 {e_img:.2e}; params rel {e_p}", flush=True)
    assert e_img <= 1e-3
    assert np.all(e_p <= 1e-3)
