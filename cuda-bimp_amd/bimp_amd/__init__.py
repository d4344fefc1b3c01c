"""bimp_amd -- Python host interface of the MI355X-native BIMP keypoint extractor.

A thin ctypes layer over the C ABI of libbimp_hip.so (include/bimp_hip.h).  It mirrors the
reference's host API (daniel-koudouna/cuda-bimp): ``CudaStuff`` (cuda_bimp.hpp:92-100),
``get_keypoints`` (cuda_bimp.hpp:107-111), ``make_lambdas_log`` (cuda_bimp.hpp:146),
``snap_transform_size`` (cuda_bimp.hpp:140) and the ``Context`` class (bimp.hpp:19-51).

There is no CPU fallback: every compute call runs the HIP kernels, and importing this module
raises if the shared library has not been built.
"""
import ctypes as C
import importlib.util
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbimp_hip.so")

NUM_ORI = 8
KP_VECTOR_LENGTH = 7
KP_MAX_POINTS = 8000

OK, W_OVERFLOW, E_ARG, E_ALLOC, E_DEVICE, E_UNSUPPORTED = 0, 1, -1, -2, -3, -4
MAP_DOUBLE, MAP_SINGLE, MAP_COMPLEX = 0, 1, 2
FLAG_PROFILE = 1
FLAG_SINGLE_STREAM = 2
FLAG_GRAPH = 4
FLAG_RESIZE_CPU = 8
PRECISION_FP32, PRECISION_BF16, PRECISION_FP16 = 0, 1, 2
DIRECT_MAX_KERNEL = 55

# SURF_CUDA keypoint-matrix rows (utils.cpp:109-150)
X_ROW, Y_ROW, LAPLACIAN_ROW, OCTAVE_ROW, SIZE_ROW, ANGLE_ROW, HESSIAN_ROW = range(7)

KEYPOINT_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                           ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])


class BimpError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"bimp status {status}: {msg}")
        self.status = status


class Options(C.Structure):
    _fields_ = [("gamma", C.c_float), ("sigma_ratio", C.c_float), ("support", C.c_float),
                ("pyr_threshold", C.c_float), ("nms_threshold", C.c_double),
                ("input_scale", C.c_float), ("flags", C.c_int), ("precision", C.c_int),
                ("direct_max_lambda", C.c_float), ("direct_sigmas", C.c_float)]


class ScaleInfo(C.Structure):
    _fields_ = [("lambda2", C.c_float), ("lam", C.c_float), ("pyrstep", C.c_float),
                ("kernel_size", C.c_int), ("kernel_centre", C.c_int), ("border", C.c_int),
                ("level", C.c_int), ("data_h", C.c_int), ("data_w", C.c_int),
                ("fft_h", C.c_int), ("fft_w", C.c_int), ("map_h", C.c_int),
                ("map_w", C.c_int), ("map_pitch", C.c_int), ("precision", C.c_int),
                ("tap_extent", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ProfileEntry(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int), ("total_ms", C.c_double),
                ("bytes", C.c_double)]


# Every symbol include/bimp_hip.h declares (checked by the tests).
EXPORTED = [
    "bimp_default_options", "bimp_snap_transform_size", "bimp_make_lambdas_log", "bimp_create",
    "bimp_destroy", "bimp_num_scales", "bimp_get_scale_info", "bimp_detect_u8",
    "bimp_detect_u8_matrix", "bimp_filter_u8", "bimp_cells", "bimp_keypoints", "bimp_get_map",
    "bimp_set_map", "bimp_keypoints_to_matrix", "bimp_matrix_to_keypoints", "bimp_last_error",
    "bimp_status_string", "bimp_get_profile", "bimp_reset_profile", "bimp_set_flags",
    "bimp_device_count", "bimp_device_malloc", "bimp_device_free", "bimp_memcpy_h2d",
    "bimp_memcpy_d2h", "bimp_stream_create", "bimp_stream_destroy", "bimp_stream_sync",
    "bimp_time_detect", "bimp_copy_kernel", "bimp_synth_frame",
    "bimp_batch_create", "bimp_batch_destroy", "bimp_batch_frame", "bimp_batch_run",
    "bimp_batch_results", "bimp_batch_outputs", "bimp_batch_time", "bimp_batch_plan",
    "bimp_batch_last_error",
    "bimp_diff_create", "bimp_diff_destroy", "bimp_diff_response_size", "bimp_diff_set_params",
    "bimp_diff_forward", "bimp_diff_backward", "bimp_diff_last_error", "bimp_bgr_to_gray",
    "bimp_keypoints_cpu_mode", "bimp_time_detect_frames", "bimp_get_level",
    "bimp_color_features", "bimp_match_knn2", "bimp_ratio_matches",
    "bimp_min_size_shape", "bimp_resize_linear_u8", "bimp_get_gpu_image", "bimp_map_view",
    "bimp_memcpy_h2d_2d", "bimp_detect_u8_noalloc", "bimp_integral_image",
    "bimp_surf_descriptors",
]


def _bind_torch_runtime():
    """One process must hold one HIP runtime.  PyTorch ships its own libamdhip64 /
    libhsa-runtime64 (same SONAMEs as /opt/rocm's); when PyTorch is installed, load those first
    so that libbimp_hip.so binds to them by SONAME and a later `import torch` finds them
    already loaded.  Device pointers and stream handles then pass freely between the two."""
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin or os.environ.get("BIMP_SYSTEM_HIP_RUNTIME"):
        return
    libdir = os.path.join(os.path.dirname(spec.origin), "lib")
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        path = os.path.join(libdir, name)
        if os.path.exists(path):
            C.CDLL(path, mode=C.RTLD_GLOBAL)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C cuda-bimp_amd` "
                          "or __graft_entry__.build()")
    _bind_torch_runtime()
    L = C.CDLL(LIB_PATH)
    P, I, SZ = C.c_void_p, C.c_int, C.c_size_t
    sig = {
        "bimp_default_options": (None, [C.POINTER(Options)]),
        "bimp_snap_transform_size": (I, [I]),
        "bimp_make_lambdas_log": (I, [C.c_float, C.c_float, I, P, I]),
        "bimp_create": (I, [I, I, P, I, C.POINTER(Options), C.POINTER(P)]),
        "bimp_destroy": (I, [P]),
        "bimp_num_scales": (I, [P]),
        "bimp_get_scale_info": (I, [P, I, C.POINTER(ScaleInfo)]),
        "bimp_detect_u8": (I, [P, P, SZ, I, P, I, I, C.POINTER(I), P]),
        "bimp_detect_u8_matrix": (I, [P, P, SZ, I, P, I, I, I, C.POINTER(I), P]),
        "bimp_filter_u8": (I, [P, P, SZ, I, P]),
        "bimp_cells": (I, [P, P]),
        "bimp_keypoints": (I, [P, P, I, I, C.POINTER(I), P]),
        "bimp_get_map": (I, [P, I, I, I, P, SZ, I, P]),
        "bimp_set_map": (I, [P, I, I, I, P, SZ, I, P]),
        "bimp_keypoints_to_matrix": (I, [P, I, P, I]),
        "bimp_matrix_to_keypoints": (I, [P, I, I, P]),
        "bimp_last_error": (C.c_char_p, [P]),
        "bimp_status_string": (C.c_char_p, [I]),
        "bimp_get_profile": (I, [P, P, I, C.POINTER(I)]),
        "bimp_reset_profile": (I, [P]),
        "bimp_set_flags": (I, [P, I]),
        "bimp_device_count": (I, [C.POINTER(I)]),
        "bimp_device_malloc": (I, [C.POINTER(P), SZ]),
        "bimp_device_free": (I, [P]),
        "bimp_memcpy_h2d": (I, [P, P, SZ]),
        "bimp_memcpy_d2h": (I, [P, P, SZ]),
        "bimp_stream_create": (I, [C.POINTER(P)]),
        "bimp_stream_destroy": (I, [P]),
        "bimp_stream_sync": (I, [P]),
        "bimp_time_detect": (I, [P, P, SZ, P, I, I, P, C.POINTER(C.c_float)]),
        "bimp_copy_kernel": (I, [P, P, SZ, P]),
        "bimp_time_detect_frames": (I, [P, P, SZ, P, I, I, P, P]),
        "bimp_get_level": (I, [P, I, P, SZ, P]),
        "bimp_color_features": (I, [P, I, I, P, SZ, I, I, P, I, P]),
        "bimp_min_size_shape": (I, [I, I, C.POINTER(I), C.POINTER(I), C.POINTER(C.c_float)]),
        "bimp_resize_linear_u8": (I, [P, SZ, I, I, I, P, SZ, I, I, C.c_double, C.c_double, P]),
        "bimp_get_gpu_image": (I, [P, SZ, I, I, I, I, P, SZ, P]),
        "bimp_memcpy_h2d_2d": (I, [P, SZ, P, SZ, SZ, SZ]),
        "bimp_detect_u8_noalloc": (I, [P, P, SZ, I, P, I, C.POINTER(I), P]),
        "bimp_integral_image": (I, [P, SZ, I, I, P, P]),
        "bimp_surf_descriptors": (I, [P, I, I, P, I, I, I, P, I, P, P]),
        "bimp_map_view": (I, [P, I, I, I, C.POINTER(P), C.POINTER(SZ), C.POINTER(I), C.POINTER(I)]),
        "bimp_match_knn2": (I, [P, I, I, P, I, I, I, P, P, P]),
        "bimp_ratio_matches": (I, [P, P, I, C.c_double, P, I, C.POINTER(I)]),
        "bimp_synth_frame": (I, [I, I, I, C.c_uint64, P, SZ]),
        "bimp_batch_create": (I, [P, P, I, P, I, C.POINTER(Options), I, I, C.POINTER(P)]),
        "bimp_batch_destroy": (I, [P]),
        "bimp_batch_frame": (I, [P, I, C.POINTER(P), C.POINTER(SZ)]),
        "bimp_batch_run": (I, [P, P]),
        "bimp_batch_results": (I, [P, I, P, I, C.POINTER(I)]),
        "bimp_batch_outputs": (I, [P, I, C.POINTER(P), C.POINTER(P)]),
        "bimp_batch_time": (I, [P, I, P, C.POINTER(C.c_float), P]),
        "bimp_batch_plan": (I, [P, C.c_char_p, I]),
        "bimp_batch_last_error": (C.c_char_p, [P]),
        "bimp_diff_create": (I, [I, I, P, I, C.POINTER(Options), C.POINTER(P)]),
        "bimp_diff_destroy": (I, [P]),
        "bimp_diff_response_size": (C.c_longlong, [P]),
        "bimp_diff_set_params": (I, [P, C.c_float, C.c_float]),
        "bimp_diff_forward": (I, [P, P, I, P, P]),
        "bimp_diff_backward": (I, [P, P, I, P, P, P, P]),
        "bimp_diff_last_error": (C.c_char_p, [P]),
        "bimp_bgr_to_gray": (I, [P, SZ, I, I, I, P, SZ, P]),
        "bimp_keypoints_cpu_mode": (I, [P, P, I, C.POINTER(I), P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


lib = _load()


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(st, ctx=None):
    if st < 0:
        msg = lib.bimp_last_error(ctx).decode()
        raise BimpError(st, msg or lib.bimp_status_string(st).decode())
    return st


def default_options():
    o = Options()
    lib.bimp_default_options(C.byref(o))
    return o


def snap_transform_size(n):
    """snapTransformSize (cuda_bimp.cu:200-219)."""
    return _check(lib.bimp_snap_transform_size(int(n)))


def make_lambdas_log(start, end, scales_per_octave):
    """makeLambdasLog (cuda_bimp.cu:927-937): float-accumulated geometric wavelengths."""
    n = _check(lib.bimp_make_lambdas_log(start, end, scales_per_octave, None, 0))
    out = np.zeros(max(n, 1), np.float32)
    lib.bimp_make_lambdas_log(start, end, scales_per_octave, _ptr(out), n)
    return out[:n]


def synth_frame(width, height, nshapes, seed):
    """Seeded synthetic uint8 frame (rectangles/ellipses on a constant level, sigma-5 noise)."""
    out = np.zeros((height, width), np.uint8)
    _check(lib.bimp_synth_frame(width, height, nshapes, seed, _ptr(out), width))
    return out


def device_count():
    n = C.c_int(0)
    lib.bimp_device_count(C.byref(n))
    return n.value


class DeviceBuffer:
    """Raw device allocation (plumbing for device-resident frames and outputs)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib.bimp_device_malloc(C.byref(p), self.nbytes))
        self.ptr = p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib.bimp_memcpy_h2d(self.ptr, _ptr(arr), arr.nbytes))

    def download(self, arr):
        assert arr.flags.c_contiguous and arr.nbytes <= self.nbytes
        _check(lib.bimp_memcpy_d2h(_ptr(arr), self.ptr, arr.nbytes))
        return arr

    def free(self):
        # (at interpreter exit the module's `lib` may already be gone; the process is ending)
        if getattr(self, "ptr", None) and lib is not None:
            lib.bimp_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _check(lib.bimp_stream_create(C.byref(s)))
        self.handle = s

    def sync(self):
        _check(lib.bimp_stream_sync(self.handle))

    def __del__(self):
        if getattr(self, "handle", None) and lib is not None:
            lib.bimp_stream_destroy(self.handle)
        self.handle = None


class CudaStuff:
    """Per-frame-size state: bank spectra and buffers (CudaStuff, cuda_bimp.cu:222-460)."""

    def __init__(self, width, height, lambdas, options=None):
        lam = np.ascontiguousarray(lambdas, np.float32)
        ctx = C.c_void_p()
        opt = options if options is not None else default_options()
        _check(lib.bimp_create(int(width), int(height), _ptr(lam), lam.size, C.byref(opt),
                               C.byref(ctx)))
        self.ctx = ctx
        self.width, self.height = int(width), int(height)
        self.options = opt
        self.ks = []
        for i in range(lib.bimp_num_scales(ctx)):
            si = ScaleInfo()
            _check(lib.bimp_get_scale_info(ctx, i, C.byref(si)), ctx)
            self.ks.append(si)

    def close(self):
        if getattr(self, "ctx", None) and lib is not None:
            lib.bimp_destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        self.close()

    def _chk(self, st):
        return _check(st, self.ctx)

    # ---- stages
    def filter(self, img, stream=None):
        img = np.ascontiguousarray(img, np.uint8)
        assert img.shape == (self.height, self.width)
        self._chk(lib.bimp_filter_u8(self.ctx, _ptr(img), self.width, 0, stream))

    def cells(self, stream=None):
        self._chk(lib.bimp_cells(self.ctx, stream))

    def keypoints(self, cap=KP_MAX_POINTS, stream=None):
        out = np.zeros(max(cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        self._chk(lib.bimp_keypoints(self.ctx, _ptr(out), cap, 0, C.byref(n), stream))
        return out[:min(n.value, cap)].copy(), n.value

    def detect(self, img, cap=KP_MAX_POINTS, stream=None):
        img = np.ascontiguousarray(img, np.uint8)
        assert img.shape == (self.height, self.width)
        out = np.zeros(max(cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        self._chk(lib.bimp_detect_u8(self.ctx, _ptr(img), self.width, 0, _ptr(out), cap, 0,
                                     C.byref(n), stream))
        return out[:min(n.value, cap)].copy(), n.value

    def detect_device(self, d_img, pitch, cap=KP_MAX_POINTS, stream=None):
        """bimp_detect_u8 on a device frame (DeviceBuffer, row pitch in bytes)."""
        out = np.zeros(max(cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        self._chk(lib.bimp_detect_u8(self.ctx, d_img.ptr, int(pitch), 1, _ptr(out), cap, 0,
                                     C.byref(n), stream))
        return out[:min(n.value, cap)].copy(), n.value

    def detect_noalloc(self, img, cap=KP_MAX_POINTS, stream=None, pitch=None):
        """bimp_detect_u8_noalloc: a host frame, or a DeviceBuffer frame with its pitch."""
        out = np.zeros(max(cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        if isinstance(img, DeviceBuffer):
            st = lib.bimp_detect_u8_noalloc(self.ctx, img.ptr, pitch or self.width, 1, _ptr(out),
                                            cap, C.byref(n), stream)
        else:
            img = np.ascontiguousarray(img, np.uint8)
            assert img.shape == (self.height, self.width)
            st = lib.bimp_detect_u8_noalloc(self.ctx, _ptr(img), self.width, 0, _ptr(out), cap,
                                            C.byref(n), stream)
        self._chk(st)
        return out[:min(n.value, cap)].copy(), n.value

    def detect_matrix(self, img, cap=KP_MAX_POINTS, stream=None):
        img = np.ascontiguousarray(img, np.uint8)
        mat = np.zeros((KP_VECTOR_LENGTH, max(cap, 1)), np.float32)
        n = C.c_int(0)
        self._chk(lib.bimp_detect_u8_matrix(self.ctx, _ptr(img), self.width, 0, _ptr(mat),
                                            mat.shape[1], cap, 0, C.byref(n), stream))
        return mat[:, :min(n.value, cap)].copy(), n.value

    def get_map(self, scale, which, orientation=0, stream=None):
        si = self.ks[scale]
        if which == MAP_COMPLEX:
            shape = (si.map_h, si.map_w)
        else:
            shape = (si.data_h, si.data_w)
        out = np.zeros(shape, np.float32)
        self._chk(lib.bimp_get_map(self.ctx, scale, which, orientation, _ptr(out), shape[1], 0,
                                   stream))
        return out

    def set_map(self, scale, which, arr, orientation=0, stream=None):
        arr = np.ascontiguousarray(arr, np.float32)
        self._chk(lib.bimp_set_map(self.ctx, scale, which, orientation, _ptr(arr),
                                   arr.shape[1], 0, stream))

    # ---- profiling
    def keypoints_cpu_mode(self, cap=KP_MAX_POINTS, stream=None):
        """The reference's CPU keypoint mode on the current D and S maps."""
        out = np.zeros(max(cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        self._chk(lib.bimp_keypoints_cpu_mode(self.ctx, _ptr(out), cap, C.byref(n), stream))
        return out[:min(n.value, cap)].copy(), n.value

    def set_flags(self, flags):
        self._chk(lib.bimp_set_flags(self.ctx, flags))

    def reset_profile(self):
        self._chk(lib.bimp_reset_profile(self.ctx))

    def profile(self):
        buf = (ProfileEntry * 32)()
        n = C.c_int(0)
        self._chk(lib.bimp_get_profile(self.ctx, buf, 32, C.byref(n)))
        return {buf[i].name.decode(): {"launches": buf[i].launches, "total_ms": buf[i].total_ms,
                                       "bytes": buf[i].bytes} for i in range(n.value)}

    def time_detect(self, d_img, pitch, d_out, cap, iters, stream):
        ms = C.c_float(0)
        self._chk(lib.bimp_time_detect(self.ctx, d_img.ptr, pitch, d_out.ptr, cap, iters,
                                       stream.handle if stream else None, C.byref(ms)))
        return ms.value

    def time_detect_frames(self, d_img, pitch, d_out, cap, iters, stream):
        """Per-frame elapsed milliseconds of `iters` back-to-back frames (numpy array)."""
        ms = np.zeros(int(iters), np.float32)
        self._chk(lib.bimp_time_detect_frames(self.ctx, d_img.ptr, pitch, d_out.ptr, cap,
                                              int(iters), stream.handle if stream else None,
                                              _ptr(ms)))
        return ms

    def get_level(self, level, stream=None):
        """Pyramid level `level` (>= 1) of the last filtered frame, uint8 (h, w)."""
        w, h = self.width, self.height
        for _ in range(level):
            w, h = (w + 1) // 2, (h + 1) // 2
        out = np.zeros((h, w), np.uint8)
        self._chk(lib.bimp_get_level(self.ctx, int(level), _ptr(out), w,
                                     stream.handle if stream else None))
        return out


def bgr_to_gray(bgr):
    """8-bit BGR (H, W, 3) -> gray (H, W) on the device (getGPUImage's cvtColor step)."""
    bgr = np.ascontiguousarray(bgr, np.uint8)
    h, w, c = bgr.shape
    assert c == 3
    d = DeviceBuffer(w * h)
    _check(lib.bimp_bgr_to_gray(_ptr(bgr), 3 * w, w, h, 0, d.ptr, w, None))
    out = np.zeros((h, w), np.uint8)
    d.download(out)
    d.free()
    return out


def min_size_shape(width, height):
    """(width, height, factor) of getGPUImage's prepared frame (cuda_bimp.cu:156-168)."""
    w, h, f = C.c_int(0), C.c_int(0), C.c_float(0)
    _check(lib.bimp_min_size_shape(int(width), int(height), C.byref(w), C.byref(h), C.byref(f)))
    return w.value, h.value, f.value


def get_gpu_image(img, grayscale=True):
    """getGPUImage on the device: the minimum-size upscale of a uint8 frame (H, W) or BGR
    (H, W, 3), then BGR -> gray when grayscale; returns the prepared host copy."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    cn = 1 if img.ndim == 2 else img.shape[2]
    ow, oh, _ = min_size_shape(w, h)
    ocn = 1 if (grayscale and cn == 3) else cn
    out = np.zeros((oh, ow) if ocn == 1 else (oh, ow, ocn), np.uint8)
    ds, dd = DeviceBuffer(img.nbytes), DeviceBuffer(out.nbytes)
    ds.upload(img)
    _check(lib.bimp_get_gpu_image(ds.ptr, w * cn, w, h, cn, 1 if grayscale else 0, dd.ptr,
                                  ow * ocn, None))
    dd.download(out)
    return out


def resize_linear(img, dw, dh, fx=0.0, fy=0.0):
    """Bilinear resize (cv::resize INTER_LINEAR semantics) of a uint8 frame on the device."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    cn = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((dh, dw) if cn == 1 else (dh, dw, cn), np.uint8)
    ds, dd = DeviceBuffer(img.nbytes), DeviceBuffer(out.nbytes)
    ds.upload(img)
    _check(lib.bimp_resize_linear_u8(ds.ptr, w * cn, w, h, cn, dd.ptr, dw * cn, dw, dh,
                                     float(fx), float(fy), None))
    dd.download(out)
    return out


COLOR_DIMENSIONS = 24
MATCH_DTYPE = np.dtype([("query_idx", "<i4"), ("train_idx", "<i4"), ("img_idx", "<i4"),
                        ("distance", "<f4")])


def color_features(kp_matrix, bgr):
    """Colour histogram (n x 24 floats) of the keypoints of a 7-row SURF_CUDA matrix over a
    BGR frame (H, W, 3), computed on the device (addColorFeatures, cuda_bimp.cuh:8-71)."""
    mat = np.ascontiguousarray(kp_matrix, np.float32)
    bgr = np.ascontiguousarray(bgr, np.uint8)
    h, w, c = bgr.shape
    assert c == 3 and mat.shape[0] == 7
    n = mat.shape[1]
    out = np.zeros((n, COLOR_DIMENSIONS), np.float32)
    if n == 0:
        return out
    dk, db, dd = DeviceBuffer(mat.nbytes), DeviceBuffer(bgr.nbytes), DeviceBuffer(out.nbytes)
    dk.upload(mat)
    db.upload(bgr)
    _check(lib.bimp_color_features(dk.ptr, n, n, db.ptr, 3 * w, w, h, dd.ptr, COLOR_DIMENSIONS,
                                   None))
    dd.download(out)
    return out


def surf_descriptors(gray, kp_matrix, upright=False):
    """SURF descriptors (n x 64) and angles (n,) of the keypoints of a 7-row matrix over a gray
    uint8 frame, on the device (integral image, then one wavefront per keypoint)."""
    gray = np.ascontiguousarray(gray, np.uint8)
    mat = np.ascontiguousarray(kp_matrix, np.float32)
    h, w = gray.shape
    n = mat.shape[1]
    desc = np.zeros((n, 64), np.float32)
    ang = np.zeros(n, np.float32)
    if n == 0:
        return desc, ang
    dg, di = DeviceBuffer(gray.nbytes), DeviceBuffer((w + 1) * (h + 1) * 4)
    dk, dd, da = DeviceBuffer(mat.nbytes), DeviceBuffer(desc.nbytes), DeviceBuffer(ang.nbytes)
    dg.upload(gray)
    dk.upload(mat)
    _check(lib.bimp_integral_image(dg.ptr, w, w, h, di.ptr, None))
    _check(lib.bimp_surf_descriptors(di.ptr, w, h, dk.ptr, n, n, 1 if upright else 0, dd.ptr, 64,
                                     da.ptr, None))
    dd.download(desc)
    da.download(ang)
    return desc, ang


def match_knn2(query, train):
    """Two nearest train rows (L2) of every query row on the device: (idx2, dist2), each
    (nq, 2) (cv::cuda BFMatcher knnMatch, k = 2)."""
    q = np.ascontiguousarray(query, np.float32)
    t = np.ascontiguousarray(train, np.float32)
    nq, dim = q.shape
    nt = t.shape[0]
    assert t.shape[1] == dim
    idx = np.zeros((nq, 2), np.int32)
    dist = np.zeros((nq, 2), np.float32)
    if nq == 0:
        return idx, dist
    dq, dt = DeviceBuffer(q.nbytes), DeviceBuffer(max(t.nbytes, 4))
    di, dd = DeviceBuffer(idx.nbytes), DeviceBuffer(dist.nbytes)
    dq.upload(q)
    if nt:
        dt.upload(t)
    _check(lib.bimp_match_knn2(dq.ptr, nq, dim, dt.ptr, nt, dim, dim, di.ptr, dd.ptr, None))
    di.download(idx)
    dd.download(dist)
    return idx, dist


def ratio_matches(idx2, dist2, ratio=0.6):
    """Matches passing the ratio test (downloadMatches, bimp.cpp:208-240), cv::DMatch fields."""
    idx2 = np.ascontiguousarray(idx2, np.int32)
    dist2 = np.ascontiguousarray(dist2, np.float32)
    nq = idx2.shape[0]
    out = np.zeros(max(nq, 1), MATCH_DTYPE)
    n = C.c_int(0)
    _check(lib.bimp_ratio_matches(_ptr(idx2), _ptr(dist2), nq, float(ratio), _ptr(out), nq,
                                  C.byref(n)))
    return out[:n.value].copy()


class Batch:
    """Batching form: frames of uneven sizes, one hipGraph launch per batch (bimp_batch_*)."""

    def __init__(self, sizes, lambdas, options=None, cap=KP_MAX_POINTS, tune=False):
        w = np.ascontiguousarray([s[0] for s in sizes], np.int32)
        h = np.ascontiguousarray([s[1] for s in sizes], np.int32)
        lam = np.ascontiguousarray(lambdas, np.float32)
        opt = options if options is not None else default_options()
        b = C.c_void_p()
        st = lib.bimp_batch_create(_ptr(w), _ptr(h), len(sizes), _ptr(lam), lam.size,
                                   C.byref(opt), int(cap), 1 if tune else 0, C.byref(b))
        if st < 0:
            raise BimpError(st, lib.bimp_status_string(st).decode())
        self.b = b
        self.sizes = [(int(a), int(c)) for a, c in sizes]
        self.cap = int(cap)

    def _chk(self, st):
        if st < 0:
            raise BimpError(st, lib.bimp_batch_last_error(self.b).decode() or
                            lib.bimp_status_string(st).decode())
        return st

    def upload(self, i, img):
        """Copy a host uint8 frame into frame slot i."""
        w, h = self.sizes[i]
        img = np.ascontiguousarray(img, np.uint8)
        assert img.shape == (h, w)
        d, pitch = C.c_void_p(), C.c_size_t()
        self._chk(lib.bimp_batch_frame(self.b, i, C.byref(d), C.byref(pitch)))
        if pitch.value == w:
            self._chk(lib.bimp_memcpy_h2d(d, _ptr(img), img.nbytes))
        else:
            padded = np.zeros((h, pitch.value), np.uint8)
            padded[:, :w] = img
            self._chk(lib.bimp_memcpy_h2d(d, _ptr(padded), padded.nbytes))

    def run(self, stream=None):
        self._chk(lib.bimp_batch_run(self.b, stream.handle if stream else None))

    def results(self, i):
        out = np.zeros(max(self.cap, 1), KEYPOINT_DTYPE)
        n = C.c_int(0)
        self._chk(lib.bimp_batch_results(self.b, i, _ptr(out), self.cap, C.byref(n)))
        return out[:min(n.value, self.cap)].copy(), n.value

    def time(self, iters, stream, per_frame=True):
        ms = C.c_float(0)
        fr = np.zeros(len(self.sizes), np.float32)
        self._chk(lib.bimp_batch_time(self.b, int(iters), stream.handle, C.byref(ms),
                                      _ptr(fr) if per_frame else None))
        return ms.value, (fr if per_frame else None)

    def plan(self):
        buf = C.create_string_buffer(4096)
        lib.bimp_batch_plan(self.b, buf, 4096)
        return buf.value.decode()

    def close(self):
        if getattr(self, "b", None) and lib is not None:
            lib.bimp_batch_destroy(self.b)
        self.b = None

    def __del__(self):
        self.close()


class Diff:
    """Differentiation form (bimp_diff_*): float images -> per-scale D maps, and the adjoint.
    Device buffers are passed as raw pointers (e.g. torch tensors' data_ptr())."""

    def __init__(self, width, height, lambdas, options=None):
        lam = np.ascontiguousarray(lambdas, np.float32)
        opt = options if options is not None else default_options()
        d = C.c_void_p()
        st = lib.bimp_diff_create(int(width), int(height), _ptr(lam), lam.size, C.byref(opt),
                                  C.byref(d))
        if st < 0:
            raise BimpError(st, lib.bimp_status_string(st).decode())
        self.d = d
        self.width, self.height = int(width), int(height)
        self.response_size = int(lib.bimp_diff_response_size(d))

    def _chk(self, st):
        if st < 0:
            raise BimpError(st, lib.bimp_diff_last_error(self.d).decode() or
                            lib.bimp_status_string(st).decode())
        return st

    def set_params(self, sigma_ratio, gamma):
        self._chk(lib.bimp_diff_set_params(self.d, float(sigma_ratio), float(gamma)))

    def forward(self, img_ptr, nimg, resp_ptr, stream=None):
        self._chk(lib.bimp_diff_forward(self.d, C.c_void_p(img_ptr), int(nimg),
                                        C.c_void_p(resp_ptr), stream))

    def backward(self, img_ptr, nimg, gresp_ptr, gimg_ptr, stream=None):
        gp = np.zeros(2, np.float64)
        self._chk(lib.bimp_diff_backward(self.d, C.c_void_p(img_ptr), int(nimg),
                                         C.c_void_p(gresp_ptr), C.c_void_p(gimg_ptr), _ptr(gp),
                                         stream))
        return gp

    def close(self):
        if getattr(self, "d", None) and lib is not None:
            lib.bimp_diff_destroy(self.d)
        self.d = None

    def __del__(self):
        self.close()


def get_keypoints(img, cs, cap=KP_MAX_POINTS):
    """bimp::cuda::getKeypoints (cuda_bimp.cu:109-140) + downloadKeypoints (utils.cpp:130-150):
    returns the keypoint records (cv::KeyPoint field order) and the total count."""
    return cs.detect(img, cap)


def keypoints_to_matrix(kps):
    """utils::uploadKeypoints layout (utils.cpp:109-128): 7 x N SURF_CUDA matrix."""
    kps = np.ascontiguousarray(kps, KEYPOINT_DTYPE)
    n = kps.size
    mat = np.zeros((KP_VECTOR_LENGTH, max(n, 1)), np.float32)
    _check(lib.bimp_keypoints_to_matrix(_ptr(kps), n, _ptr(mat), mat.shape[1]))
    return mat[:, :n]


def matrix_to_keypoints(mat):
    """utils::downloadKeypoints layout (utils.cpp:130-150)."""
    mat = np.ascontiguousarray(mat, np.float32)
    n = mat.shape[1]
    kps = np.zeros(max(n, 1), KEYPOINT_DTYPE)
    _check(lib.bimp_matrix_to_keypoints(_ptr(mat), mat.shape[1], n, _ptr(kps)))
    return kps[:n]


class Context:
    """bimp::Context (bimp.hpp:19-51, bimp.cpp:7-129) for grayscale uint8 frames.

    The default bank is makeLambdasLog(4, 64, 2) as in the reference constructor
    (bimp.cpp:8); ``lambdas`` overrides it (the reference's CudaStuff takes any bank)."""

    def __init__(self, image, lambdas=None, options=None):
        image = np.ascontiguousarray(image, np.uint8)
        if image.ndim != 2:
            raise ValueError("grayscale uint8 frames only")
        self.lambdas = (make_lambdas_log(4, 64, 2) if lambdas is None
                        else np.asarray(lambdas, np.float32))
        self.cuda_data = CudaStuff(image.shape[1], image.shape[0], self.lambdas, options)
        self.current_image = image
        self.has_changed_keypoints = True
        self._keypoints = None
        self._count = 0

    def width(self):
        return self.current_image.shape[1]

    def height(self):
        return self.current_image.shape[0]

    def loadImage(self, image):
        image = np.ascontiguousarray(image, np.uint8)
        if image.shape != self.current_image.shape:
            raise ValueError("frame size differs from the context's")
        self.current_image = image
        self.has_changed_keypoints = True

    def getKeypoints(self):
        """7 x N SURF_CUDA keypoint matrix (the GpuMat the reference returns, host copy)."""
        if self.has_changed_keypoints:
            self._keypoints, self._count = self.cuda_data.detect(self.current_image)
            self.has_changed_keypoints = False
        return keypoints_to_matrix(self._keypoints)

    def getKeypointsAndDownload(self):
        if self.has_changed_keypoints:
            self.getKeypoints()
        return self._keypoints.copy()

    def getResponses(self):
        """Per-scale parameters and response maps (Context::getResponses, bimp.cpp:94-104)."""
        out = []
        for i, si in enumerate(self.cuda_data.ks):
            out.append({"info": si.as_dict(),
                        "double": self.cuda_data.get_map(i, MAP_DOUBLE),
                        "single": self.cuda_data.get_map(i, MAP_SINGLE)})
        return out

    def setCPUMode(self, cpu_mode):
        if cpu_mode:
            raise NotImplementedError("the CPU-NMS mode of the reference is not provided")
