"""The C++ Context facade (include/bimp/bimp.hpp) on the GPU: a compiled caller gets the
same keypoints as the Python binding of the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from kpmatch import bits

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("bank", ["base", "default"])
def test_context_facade_matches_c_abi(tmp_path, bank):
    """bank "default": the Context's own makeLambdasLog(4, 64, 2) (9 scales, pyramid level 4,
    the 7.9999995 wavelength); "base": the 8-wavelength bank 4 * 2^(k/4)."""
    import bimp_amd as B
    if B.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    exe = tmp_path / "facade_main"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "facade_main.cpp"),
                           "-L", os.path.dirname(B.LIB_PATH), "-lbimp_hip",
                           "-Wl,-rpath," + os.path.dirname(B.LIB_PATH), "-o", str(exe)])
    w, h = 320, 240
    img = O.synth_frame(w, h, 25, 31)
    frame = tmp_path / "frame.u8"
    img.tofile(frame)
    out = tmp_path / "out.bin"
    args = [str(exe), str(w), str(h), str(frame), str(out)] + (["default"] if bank == "default" else [])
    subprocess.check_call(args, timeout=120)
    raw = np.fromfile(out, np.uint8)
    n = int(raw[:4].view(np.int32)[0])
    kps = raw[4:4 + 28 * n].view(B.KEYPOINT_DTYPE)
    o = 4 + 28 * n
    mat = raw[o:o + 28 * n].view(np.float32).reshape(7, n)
    o += 28 * n
    nc = int(raw[o:o + 4].view(np.int32)[0])
    cpu = raw[o + 4:o + 4 + 28 * nc].view(B.KEYPOINT_DTYPE)
    lam = B.make_lambdas_log(4, 64, 2) if bank == "default" else O.BASE_BANK
    cs = B.CudaStuff(w, h, lam)
    assert len(cs.ks) == (9 if bank == "default" else 8)
    want, m = cs.detect(img)
    assert n == m
    assert np.array_equal(bits(kps), bits(want))
    assert np.array_equal(bits(B.matrix_to_keypoints(mat)), bits(want))
    want_cpu, mc = cs.keypoints_cpu_mode()
    assert nc == mc and nc > 0
    assert np.array_equal(bits(cpu), bits(want_cpu))
