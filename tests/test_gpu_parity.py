"""GPU parity: the HIP library (called through its C ABI) against the CPU oracle.

Bars: complex-cell and response maps within 1e-4 relative (max|a-b|/max|b|); keypoint
agreement >= 99% (same SIZE bin, <= 1 px); stencil and compaction stages bit-exact when fed
the oracle's own input maps.
"""
import numpy as np
import pytest

import oracle_lib as O
from kpmatch import agreement, bits, rel_err

pytestmark = pytest.mark.gpu

MAP_TOL = 1e-4


@pytest.fixture(scope="module")
def B():
    import bimp_amd
    if bimp_amd.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    return bimp_amd


def _frame(w, h, n, seed):
    return O.synth_frame(w, h, n, seed)


def _pair(B, w, h, lambdas, fp64=True):
    return B.CudaStuff(w, h, lambdas), O.Oracle(w, h, lambdas, fp64=fp64)


@pytest.mark.parametrize("w,h,bank", [
    (96, 80, 5),      # FFT 128 x 128 (H wraps)
    (160, 120, 8),    # 256 x 128, half-res 128 x 64
    (200, 150, 8),    # 256 x 256 / 128 x 128
])
def test_complex_maps_match_oracle(B, w, h, bank):
    lam = O.BASE_BANK[:bank]
    cs, orc = _pair(B, w, h, lam)
    img = _frame(w, h, 12, 7)
    orc.set_image(img)
    cs.filter(img)
    for i, si in enumerate(cs.ks):
        s = orc.scales[i]
        assert (si.data_h, si.data_w, si.fft_h, si.fft_w) == (s.dataH, s.dataW, s.fftH, s.fftW)
        assert si.kernel_size == s.kW and si.border == s.off
        ref = orc.filter(i)
        for o in range(8):
            got = cs.get_map(i, B.MAP_COMPLEX, o)
            want = ref[o, :si.map_h, :si.map_w]
            e = rel_err(got, want)
            print(f"scale {i} o {o} rel_err {e:.2e}", flush=True)
            assert e < MAP_TOL, (i, o, e)


@pytest.mark.parametrize("w,h", [(96, 80), (160, 120), (256, 256), (320, 240)])
def test_cells_bit_exact_on_oracle_maps(B, w, h):
    """Given identical complex maps, end-stopping + inhibition are bit-exact."""
    lam = O.BASE_BANK
    cs, orc = _pair(B, w, h, lam)
    img = _frame(w, h, 15, 11)
    orc.set_image(img)
    for i, si in enumerate(cs.ks):
        maps = orc.filter(i)
        for o in range(8):
            cs.set_map(i, B.MAP_COMPLEX, maps[o, :si.map_h, :si.map_w], o)
        D, S = orc.cells(i, maps)
        cs.cells()
        gD = cs.get_map(i, B.MAP_DOUBLE)
        gS = cs.get_map(i, B.MAP_SINGLE)
        assert np.array_equal(gD.view(np.uint32), D.view(np.uint32)), i
        assert np.array_equal(gS.view(np.uint32), S.view(np.uint32)), i


@pytest.mark.parametrize("w,h", [(160, 120), (320, 240)])
def test_compaction_bit_exact_on_oracle_maps(B, w, h):
    """Given identical D maps, NMS + compaction are bit-exact, in (scale, row, col) order."""
    lam = O.BASE_BANK
    cs, orc = _pair(B, w, h, lam)
    img = _frame(w, h, 20, 3)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    for i in range(len(cs.ks)):
        cs.set_map(i, B.MAP_DOUBLE, Ds[i])
    got, gn = cs.keypoints()
    assert gn == n
    assert np.array_equal(bits(got), bits(kps))


@pytest.mark.parametrize("w,h,nshapes,seed", [(96, 80, 6, 1), (200, 150, 15, 2),
                                              (256, 256, 20, 4), (320, 240, 25, 5)])
def test_detect_end_to_end(B, w, h, nshapes, seed):
    lam = O.BASE_BANK
    cs, orc = _pair(B, w, h, lam)
    img = _frame(w, h, nshapes, seed)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    got, gn = cs.detect(img)
    for i in range(len(cs.ks)):
        eD = rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i])
        eS = rel_err(cs.get_map(i, B.MAP_SINGLE), Ss[i])
        print(f"scale {i} D {eD:.2e} S {eS:.2e}", flush=True)
        assert eD < MAP_TOL and eS < MAP_TOL
    a = agreement(got, kps)
    print(f"keypoints gpu {gn} oracle {n} agreement {a:.4f}", flush=True)
    assert a >= 0.99


def test_matrix_layout_matches_keypoints(B):
    w, h = 160, 120
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = _frame(w, h, 10, 9)
    kps, n = cs.detect(img)
    mat, n2 = cs.detect_matrix(img)
    assert n == n2
    back = B.matrix_to_keypoints(mat)
    assert np.array_equal(bits(back), bits(kps))
    assert np.all(mat[B.LAPLACIAN_ROW].view(np.int32) == 1)


def test_overflow_truncates_deterministically(B):
    w, h = 256, 256
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = _frame(w, h, 40, 21)
    full, n = cs.detect(img)
    assert n > 20
    cap = n // 3
    part, n2 = cs.detect(img, cap=cap)  # BIMP_W_OVERFLOW is a warning, not an error
    assert n2 == n
    assert np.array_equal(bits(part), bits(full[:cap]))


def test_wide_frame_radix5(B):
    """FW = 2560 (radix 5) and a narrow FH."""
    w, h = 2100, 64
    lam = O.BASE_BANK[:2]
    cs, orc = _pair(B, w, h, lam)
    assert cs.ks[0].fft_w == 2560
    img = _frame(w, h, 30, 13)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    got, gn = cs.detect(img)
    for i in range(len(cs.ks)):
        assert rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i]) < MAP_TOL
    assert agreement(got, kps) >= 0.99


def test_tall_frame_radix3(B):
    """FH = 1536 (radix 3 column transforms)."""
    w, h = 48, 1100
    lam = O.BASE_BANK[:2]
    cs, orc = _pair(B, w, h, lam)
    assert cs.ks[0].fft_h == 1536
    img = _frame(w, h, 30, 17)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    got, gn = cs.detect(img)
    for i in range(len(cs.ks)):
        assert rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i]) < MAP_TOL
    assert agreement(got, kps) >= 0.99


def test_cfg0_640x480(B):
    """The CPU-runnable workload: 640x480, 40 shapes, seed 0, full bank."""
    w, h = 640, 480
    cs, orc = _pair(B, w, h, O.BASE_BANK)
    img = _frame(w, h, 40, 0)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    got, gn = cs.detect(img)
    errs = [rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i]) for i in range(8)]
    print("D rel errs", ["%.1e" % e for e in errs], flush=True)
    assert max(errs) < MAP_TOL
    a = agreement(got, kps)
    print(f"keypoints gpu {gn} oracle {n} agreement {a:.4f}", flush=True)
    assert a >= 0.99


def test_cfg1_1920x1080(B):
    """The headline workload at its full size: 1920x1080, 300 shapes, seed 1, full bank; every
    keypoint (no 8000 cap) against the FP64 oracle."""
    import os
    O.lib().bo_set_threads(min(16, os.cpu_count() or 1))
    w, h = 1920, 1080
    cs, orc = _pair(B, w, h, O.BASE_BANK)
    img = B.synth_frame(w, h, 300, 1)
    kps, n, Ds, Ss = orc.detect(img, cap=1 << 16, want_maps=True)
    got, gn = cs.detect(img, cap=1 << 16)
    errs = [rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i]) for i in range(8)]
    print("D rel errs", ["%.1e" % e for e in errs], flush=True)
    assert max(errs) < MAP_TOL
    a = agreement(got, kps)
    print(f"keypoints gpu {gn} oracle {n} agreement {a:.4f}", flush=True)
    assert gn > 8000 and a >= 0.99


def test_repeat_frames_identical(B):
    w, h = 320, 240
    cs = B.CudaStuff(w, h, O.BASE_BANK)
    img = _frame(w, h, 25, 5)
    a, n1 = cs.detect(img)
    b, n2 = cs.detect(img)
    assert n1 == n2 and np.array_equal(bits(a), bits(b))


@pytest.mark.parametrize("w,h,fw,fh", [(2600, 200, 3072, 256), (3100, 96, 3584, 128),
                                       (48, 2600, 64, 3072), (40, 3100, 64, 3584),
                                       (3700, 72, 4096, 128), (56, 3700, 64, 4096)])
def test_transform_sizes_3072_3584(B, w, h, fw, fh):
    """Sides that snap to 3072 (48*8*8, 48 = 16*3), 3584 (56*8*8, 56 = 8*7) and 4096 (two
    wavefronts per transform) in both directions."""
    lam = O.BASE_BANK[:2]
    cs, orc = _pair(B, w, h, lam)
    assert (cs.ks[0].fft_w, cs.ks[0].fft_h) == (fw, fh)
    img = _frame(w, h, 30, 19)
    kps, n, Ds, Ss = orc.detect(img, want_maps=True)
    got, gn = cs.detect(img)
    for i in range(len(cs.ks)):
        e = rel_err(cs.get_map(i, B.MAP_DOUBLE), Ds[i])
        print(f"{w}x{h} scale {i} D rel err {e:.2e}", flush=True)
        assert e < MAP_TOL
    assert agreement(got, kps) >= 0.99
