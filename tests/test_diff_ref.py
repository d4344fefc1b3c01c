"""CPU pins of the FP64 reference gradient (tests/diff_ref.py) by FP64 central differences."""
import numpy as np
import pytest

import oracle_lib as O
from diff_ref import DiffRef


def _case(w=48, h=40, lambdas=(4.0, 5.656854, 8.0), seed=3):
    img = O.synth_frame(w, h, 6, seed).astype(np.float64) / 255.0
    ref = DiffRef(w, h, lambdas)
    targets = ref.forward(img, gamma=0.6)
    return ref, img, targets


def test_forward_matches_the_uint8_oracle_at_level_0():
    """On u8/255 the float forward's level-0 D maps are the oracle's (FP64 filter)."""
    w, h = 48, 40
    img8 = O.synth_frame(w, h, 6, 3)
    ref = DiffRef(w, h, (4.0, 5.656854))
    D = ref.forward(img8.astype(np.float64) / 255.0)
    orc = O.Oracle(w, h, [4.0, 5.656854])
    _, _, Ds, _ = orc.detect(img8, want_maps=True)
    for a, b in zip(D, Ds):
        assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


@pytest.mark.parametrize("w,h", [(48, 40), (49, 41)])
def test_image_gradient_matches_central_differences(w, h):
    """(49, 41): the level-1 data (24 x 20) is the top-left crop of a 25 x 21 level."""
    ref, img, targets = _case(w, h)
    loss, g, _ = ref.gradient(img, targets)
    rng = np.random.default_rng(0)
    for _ in range(3):
        u = rng.standard_normal(img.shape)
        h = 1e-7  # small enough that no relu branch flips between the two evaluations
        fd = (ref.loss(img + h * u, targets) - ref.loss(img - h * u, targets)) / (2 * h)
        an = float((g * u).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-12), (fd, an)


@pytest.mark.parametrize("which", [0, 1])
def test_parameter_gradient_matches_central_differences(which):
    ref, img, targets = _case()
    _, _, gth = ref.gradient(img, targets)
    h = 1e-7
    p = [0.56, 0.5]
    pp, pm = list(p), list(p)
    pp[which] += h
    pm[which] -= h
    fd = (ref.loss(img, targets, *pp) - ref.loss(img, targets, *pm)) / (2 * h)
    assert abs(fd - gth[which]) <= 1e-6 * abs(fd), (fd, gth[which])
