"""The C ABI boundary, on the CPU: the library loads, exports every declared symbol, and its
host-side functions agree with the oracle; device calls fail with status codes (no exit)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import bimp_amd as B
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "bimp_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bimp_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    names = _declared()
    assert len(names) >= 30
    lib = C.CDLL(B.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(B.EXPORTED) == names


def test_host_functions_match_oracle():
    for n in [16, 17, 100, 480, 540, 1080, 1920, 2160, 3840, 1025]:
        assert B.snap_transform_size(n) == O.snap(n)
    assert np.array_equal(B.make_lambdas_log(4, 64, 2), O.make_lambdas_log(4, 64, 2))
    assert np.array_equal(B.make_lambdas_log(2, 40, 3), O.make_lambdas_log(2, 40, 3))


def test_synth_frames_identical_in_both_libraries():
    for args in [(640, 480, 40, 0), (97, 61, 7, 12345)]:
        assert np.array_equal(B.synth_frame(*args), O.synth_frame(*args))


def test_keypoint_matrix_roundtrip():
    rng = np.random.default_rng(0)
    k = np.zeros(37, B.KEYPOINT_DTYPE)
    for f in ("x", "y", "size", "response"):
        k[f] = rng.standard_normal(37).astype(np.float32)
    k["angle"] = -1
    k["octave"] = rng.integers(0, 64, 37)
    k["class_id"] = 1
    mat = B.keypoints_to_matrix(k)
    assert mat.shape == (7, 37)
    assert np.all(mat[B.LAPLACIAN_ROW].view(np.int32) == 1)
    assert np.array_equal(mat[B.OCTAVE_ROW].view(np.int32), k["octave"])
    assert np.array_equal(mat[B.HESSIAN_ROW], k["response"])
    back = B.matrix_to_keypoints(mat)
    assert np.array_equal(back.view(np.uint32), k.view(np.uint32))


def test_errors_are_status_codes():
    opt = B.default_options()
    assert abs(opt.nms_threshold - 0.104) < 1e-15 and opt.gamma == np.float32(0.5)
    ctx = C.c_void_p()
    lam = np.array([4.0], np.float32)
    # invalid arguments are rejected before any device work
    assert B.lib.bimp_create(2, 2, B._ptr(lam), 1, None, C.byref(ctx)) == B.E_ARG
    assert B.lib.bimp_last_error(None)
    bad = np.array([0.5], np.float32)  # lambda < 1
    assert B.lib.bimp_create(64, 64, B._ptr(bad), 1, None, C.byref(ctx)) == B.E_ARG
    assert B.lib.bimp_destroy(None) == B.E_ARG
    assert B.lib.bimp_status_string(B.W_OVERFLOW) == b"keypoint overflow"
    if B.device_count() == 0:
        # no GPU in this container: creation reports a device error instead of exiting
        with pytest.raises(B.BimpError) as e:
            B.CudaStuff(64, 48, [4.0])
        assert e.value.status == B.E_DEVICE
        with pytest.raises(B.BimpError) as e:
            B.Batch([(64, 48), (32, 32)], [4.0])
        assert e.value.status == B.E_DEVICE
    # precision options are validated before any device work
    opt.precision = 7
    assert B.lib.bimp_create(64, 64, B._ptr(lam), 1, C.byref(opt), C.byref(ctx)) == B.E_ARG
    opt.precision = B.PRECISION_BF16
    opt.direct_sigmas = -1.0
    assert B.lib.bimp_create(64, 64, B._ptr(lam), 1, C.byref(opt), C.byref(ctx)) == B.E_ARG
    # batch arguments
    b = C.c_void_p()
    w = np.array([64], np.int32)
    assert B.lib.bimp_batch_create(B._ptr(w), B._ptr(w), 0, B._ptr(lam), 1, None, 10, 0,
                                   C.byref(b)) == B.E_ARG
    assert B.lib.bimp_batch_destroy(None) == B.E_ARG
    # the plan search is switched off: asking for it is refused before any device work
    w2 = np.array([64, 32], np.int32)
    assert B.lib.bimp_batch_create(B._ptr(w2), B._ptr(w2), 2, B._ptr(lam), 1, None, 10, 1,
                                   C.byref(b)) == B.E_UNSUPPORTED
    assert not b.value


def test_facade_header_compiles(tmp_path):
    exe = tmp_path / "facade_main"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror",
                           "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "tests", "cpp", "facade_main.cpp"),
                           "-L", os.path.dirname(B.LIB_PATH), "-lbimp_hip",
                           "-Wl,-rpath," + os.path.dirname(B.LIB_PATH), "-o", str(exe)])
    assert exe.exists()
