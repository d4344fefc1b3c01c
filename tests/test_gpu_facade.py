"""The C++ Context facade (include/bimp/bimp.hpp) on the GPU: a compiled caller gets the
same keypoints, device keypoint matrix, response views, CPU-mode keypoints and colour
features as the Python binding of the same C ABI, including getGPUImage's minimum-size
upscale (the detector reads the top-left width x height of the upscaled frame) and BGR input.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from kpmatch import bits

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    import bimp_amd as B
    if B.device_count() < 1:
        pytest.fail("no GPU visible to the HIP library")
    path = tmp_path_factory.mktemp("facade") / "facade_main"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "facade_main.cpp"),
                           "-L", os.path.dirname(B.LIB_PATH), "-lbimp_hip",
                           "-Wl,-rpath," + os.path.dirname(B.LIB_PATH), "-o", str(path)])
    return path


class _Reader:
    def __init__(self, raw):
        self.raw, self.o = raw, 0

    def take(self, n, dtype):
        dt = np.dtype(dtype)
        v = self.raw[self.o:self.o + n * dt.itemsize].view(dt)
        self.o += n * dt.itemsize
        return v

    def int(self):
        return int(self.take(1, np.int32)[0])


@pytest.mark.parametrize("w,h,cn,bank", [
    (320, 240, 1, "base"),     # upscaled to 350 x 262, detection on the top-left 320 x 240
    (640, 480, 1, "default"),  # the Context's own bank, no upscale
    (400, 300, 3, "base"),     # BGR, rows < 350 <= cols: 467 x 350, gray on the device
])
def test_context_facade_matches_c_abi(tmp_path, exe, w, h, cn, bank):
    import bimp_amd as B
    if cn == 1:
        img = O.synth_frame(w, h, 25, 31)
    else:
        img = np.stack([O.synth_frame(w, h, 25, 31 + c) for c in range(3)], axis=-1)
    frame = tmp_path / "frame.u8"
    img.tofile(frame)
    out = tmp_path / "out.bin"
    args = [str(exe), str(w), str(h), str(cn), str(frame), str(out)]
    if bank == "default":
        args.append("default")
    subprocess.check_call(args, timeout=120)
    r = _Reader(np.fromfile(out, np.uint8))
    fw, fh = r.int(), r.int()
    pw, ph, _ = B.min_size_shape(w, h)
    assert (fw, fh) == (pw, ph)
    prepared = B.get_gpu_image(img)  # gray, pw x ph
    assert prepared.shape == (ph, pw)
    lam = B.make_lambdas_log(4, 64, 2) if bank == "default" else O.BASE_BANK
    cs = B.CudaStuff(w, h, lam)
    assert len(cs.ks) == (9 if bank == "default" else 8)
    want, m = cs.detect(np.ascontiguousarray(prepared[:h, :w]))
    n = r.int()
    kps = r.take(n, B.KEYPOINT_DTYPE)
    mat = r.take(7 * n, np.float32).reshape(7, n)
    assert n == m
    assert np.array_equal(bits(kps), bits(want))
    assert np.array_equal(bits(B.matrix_to_keypoints(mat)), bits(want))
    nv = r.int()
    assert nv == len(cs.ks)
    for i, si in enumerate(cs.ks):
        d = r.take(si.data_h * si.data_w, np.float32).reshape(si.data_h, si.data_w)
        assert np.array_equal(d.view(np.uint32), cs.get_map(i, B.MAP_DOUBLE).view(np.uint32)), i
    nc = r.int()
    cpu = r.take(nc, B.KEYPOINT_DTYPE)
    want_cpu, mc = cs.keypoints_cpu_mode()
    assert nc == mc and nc > 0
    assert np.array_equal(bits(cpu), bits(want_cpu))
    nr = r.int()
    cpu_r = r.take(nr, B.KEYPOINT_DTYPE)
    cs.set_flags(B.FLAG_RESIZE_CPU)  # the host-rounded pyramid of Context::resize_cpu
    cs.detect(np.ascontiguousarray(prepared[:h, :w]))
    want_r, mr = cs.keypoints_cpu_mode()
    cs.set_flags(0)
    assert nr == mr and np.array_equal(bits(cpu_r), bits(want_r))
    desc = r.take(n * 152, np.float32).reshape(n, 152)
    surf, _ = B.surf_descriptors(prepared, B.keypoints_to_matrix(want))
    assert np.array_equal(desc[:, :64].view(np.uint32), surf.view(np.uint32))
    assert np.all(desc[:, 64:128] == 0)
    if cn == 3:
        color = B.get_gpu_image(img, grayscale=False)
        want_col = O.color_features(B.keypoints_to_matrix(want), color)
        assert np.array_equal(desc[:, 128:].view(np.uint32), want_col.view(np.uint32))
    else:
        assert np.all(desc[:, 128:] == 0)
    nm = r.int()
    m = r.take(nm, B.MATCH_DTYPE)
    # matched against itself: the nearest neighbour is the keypoint itself at distance 0 (the
    # ratio test keeps it unless a second keypoint has the same descriptor)
    assert nm > 0.9 * n and np.all(m["query_idx"] == m["train_idx"]) and np.all(m["distance"] == 0)
    assert r.o == len(r.raw)
