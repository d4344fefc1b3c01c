"""ctypes binding of the CPU oracle (oracle/_build/libbimp_oracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or shipped.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libbimp_oracle.so")


class Keypoint(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("size", C.c_float),
                ("angle", C.c_float), ("response", C.c_float),
                ("octave", C.c_int), ("class_id", C.c_int)]


KP_DTYPE = np.dtype([("x", "<f4"), ("y", "<f4"), ("size", "<f4"), ("angle", "<f4"),
                     ("response", "<f4"), ("octave", "<i4"), ("class_id", "<i4")])


class Scale(C.Structure):
    _fields_ = [("lambda2", C.c_float), ("lam", C.c_float), ("pyrstep", C.c_float),
                ("fs", C.c_float), ("kW", C.c_int), ("kX", C.c_int), ("off", C.c_int),
                ("level", C.c_int), ("dataH", C.c_int), ("dataW", C.c_int),
                ("fftH", C.c_int), ("fftW", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.bo_snap_transform_size.argtypes = [C.c_int]
        L.bo_make_lambdas_log.argtypes = [C.c_float, C.c_float, C.c_int, P, C.c_int]
        L.bo_scale_params.argtypes = [C.c_float, C.c_int, C.c_int, C.POINTER(Scale)]
        L.bo_gabor.argtypes = [C.POINTER(Scale), C.c_int, P, P]
        L.bo_gabor.restype = None
        L.bo_pyr_down.argtypes = [P, C.c_int, C.c_int, C.c_size_t, P]
        L.bo_pyr_down_host.argtypes = [P, C.c_int, C.c_int, C.c_size_t, P]
        L.bo_set_resize_cpu.argtypes = [P, C.c_int]
        L.bo_pad_clamp.argtypes = [P, C.c_size_t, C.POINTER(Scale), P]
        L.bo_pad_clamp.restype = None
        L.bo_end_stop_offsets.argtypes = [C.c_float, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.bo_inhibition_offsets.argtypes = [C.c_float, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.bo_fft_c2c.argtypes = [P, C.c_int, C.c_int]
        L.bo_brute_filter.argtypes = [P, C.c_size_t, C.POINTER(Scale), C.c_int, P, P]
        L.bo_brute_filter.restype = None
        L.bo_create.argtypes = [C.c_int, C.c_int, P, C.c_int, C.c_int]
        L.bo_create.restype = P
        L.bo_destroy.argtypes = [P]
        L.bo_destroy.restype = None
        L.bo_num_scales.argtypes = [P]
        L.bo_get_scale.argtypes = [P, C.c_int, C.POINTER(Scale)]
        L.bo_set_threads.argtypes = [C.c_int]
        L.bo_set_image.argtypes = [P, P, C.c_size_t]
        L.bo_get_level_image.argtypes = [P, C.c_int, P]
        L.bo_filter_scale.argtypes = [P, C.c_int, P]
        L.bo_cells_scale.argtypes = [P, C.c_int, P, P, P]
        L.bo_nms_scale.argtypes = [P, C.c_int, P, P, C.c_int, C.POINTER(C.c_int)]
        L.bo_detect.argtypes = [P, P, C.c_size_t, P, C.c_int, C.POINTER(C.c_int), P, P]
        L.bo_nms_cpu_mode.argtypes = [P, C.c_int, P, P, P, C.c_int, C.POINTER(C.c_int)]
        L.bo_direct_taps.argtypes = [C.POINTER(Scale), C.c_float, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_double), P, P]
        L.bo_direct_filter_scale.argtypes = [P, C.c_int, C.c_float, C.c_int, P]
        L.bimp_synth_frame.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, P, C.c_size_t]
        L.bo_color_features.argtypes = [P, C.c_int, C.c_int, P, C.c_size_t, C.c_int, C.c_int, P,
                                        C.c_int]
        L.bo_resize_linear.argtypes = [P, C.c_size_t, C.c_int, C.c_int, C.c_int, P, C.c_size_t,
                                       C.c_int, C.c_int, C.c_double, C.c_double]
        L.bo_integral_image.argtypes = [P, C.c_size_t, C.c_int, C.c_int, P]
        L.bo_surf_descriptors.argtypes = [P, C.c_size_t, C.c_int, C.c_int, P, C.c_int, C.c_int,
                                          C.c_int, P, C.c_int, P]
        L.bo_knn2.argtypes = [P, C.c_int, C.c_int, P, C.c_int, C.c_int, C.c_int, P, P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def synth_frame(w, h, nshapes, seed):
    out = np.zeros((h, w), np.uint8)
    assert lib().bimp_synth_frame(w, h, nshapes, seed, ptr(out), w) == 0
    return out


def color_features(kp_matrix, bgr):
    """Oracle colour histogram: (n, 24) floats for a 7 x n keypoint matrix and a BGR frame."""
    mat = np.ascontiguousarray(kp_matrix, np.float32)
    bgr = np.ascontiguousarray(bgr, np.uint8)
    h, w, _ = bgr.shape
    n = mat.shape[1]
    out = np.zeros((n, 24), np.float32)
    assert lib().bo_color_features(ptr(mat), n, n, ptr(bgr), 3 * w, w, h, ptr(out), 24) == 0
    return out


def resize_linear(img, dw, dh, fx, fy):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    cn = 1 if img.ndim == 2 else img.shape[2]
    out = np.zeros((dh, dw) if cn == 1 else (dh, dw, cn), np.uint8)
    assert lib().bo_resize_linear(ptr(img), w * cn, w, h, cn, ptr(out), dw * cn, dw, dh, fx, fy) == 0
    return out


def integral_image(img):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    out = np.zeros((h + 1, w + 1), np.int64)
    assert lib().bo_integral_image(ptr(img), w, w, h, ptr(out)) == 0
    return out


def surf_descriptors(img, kp_matrix, upright=False):
    """(n, 64) SURF descriptors and (n,) angles of the keypoints of a 7 x n matrix."""
    img = np.ascontiguousarray(img, np.uint8)
    mat = np.ascontiguousarray(kp_matrix, np.float32)
    h, w = img.shape
    n = mat.shape[1]
    desc = np.zeros((n, 64), np.float32)
    ang = np.zeros(n, np.float64)
    assert lib().bo_surf_descriptors(ptr(img), w, w, h, ptr(mat), n, n, int(upright), ptr(desc),
                                     64, ptr(ang)) == 0
    return desc, ang


def knn2(query, train):
    q = np.ascontiguousarray(query, np.float32)
    t = np.ascontiguousarray(train, np.float32)
    nq, dim = q.shape
    idx = np.zeros((nq, 2), np.int32)
    dist = np.zeros((nq, 2), np.float32)
    assert lib().bo_knn2(ptr(q), nq, dim, ptr(t), t.shape[0], dim, dim, ptr(idx), ptr(dist)) == 0
    return idx, dist


def snap(n):
    return lib().bo_snap_transform_size(n)


def make_lambdas_log(start, end, spo):
    out = np.zeros(1024, np.float32)
    n = lib().bo_make_lambdas_log(start, end, spo, ptr(out), 1024)
    return out[:n].copy()


def scale_params(lambda2, w, h):
    s = Scale()
    assert lib().bo_scale_params(lambda2, w, h, C.byref(s)) == 0
    return s


def gabor(s, o):
    ke = np.zeros((s.kW, s.kW), np.float32)
    ko = np.zeros((s.kW, s.kW), np.float32)
    lib().bo_gabor(C.byref(s), o, ptr(ke), ptr(ko))
    return ke, ko


def direct_taps(s, nsig=0.0, fp16=False):
    """(ext, fscale, dc[16], taps[16, kW, kW]) of the precision form."""
    ext = C.c_int()
    fscale = C.c_double()
    dc = np.zeros(16, np.float64)
    taps = np.zeros((16, s.kW, s.kW), np.float64)
    assert lib().bo_direct_taps(C.byref(s), nsig, int(fp16), C.byref(ext), C.byref(fscale),
                                ptr(dc), ptr(taps)) == 0
    return ext.value, fscale.value, dc, taps


def pyr_down(img, half_up=False):
    """One pyramid level: cv::cuda::pyrDown rounding (half to even), or with half_up the host
    cv::pyrDown rounding of the reference's resize_cpu option."""
    h, w = img.shape
    img = np.ascontiguousarray(img)
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), np.uint8)
    fn = lib().bo_pyr_down_host if half_up else lib().bo_pyr_down
    assert fn(ptr(img), w, h, w, ptr(out)) == 0
    return out


def pad_clamp(img, s):
    img = np.ascontiguousarray(img)
    out = np.zeros((s.fftH, s.fftW), np.float32)
    lib().bo_pad_clamp(ptr(img), img.shape[1], C.byref(s), ptr(out))
    return out


def end_stop_offsets(lam, o):
    a, b = C.c_int(), C.c_int()
    lib().bo_end_stop_offsets(lam, o, C.byref(a), C.byref(b))
    return a.value, b.value


def inhibition_offsets(lam, o):
    a, b = C.c_int(), C.c_int()
    lib().bo_inhibition_offsets(lam, o, C.byref(a), C.byref(b))
    return a.value, b.value


def fft_c2c(x, sign=-1):
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    assert lib().bo_fft_c2c(ptr(a), a.size, sign) == 0
    return a


def brute_filter(level_img, s, o):
    level_img = np.ascontiguousarray(level_img)
    e = np.zeros((s.fftH, s.fftW), np.float64)
    od = np.zeros((s.fftH, s.fftW), np.float64)
    lib().bo_brute_filter(ptr(level_img), level_img.shape[1], C.byref(s), o, ptr(e), ptr(od))
    return e, od


class Oracle:
    """Context: bank built at creation (as CudaStuff does); stages callable one by one."""

    def __init__(self, w, h, lambdas, fp64=True):
        self.w, self.h = w, h
        lam = np.ascontiguousarray(lambdas, np.float32)
        self.c = lib().bo_create(w, h, ptr(lam), lam.size, 1 if fp64 else 0)
        if not self.c:
            raise ValueError("oracle rejected the configuration")
        self.scales = []
        for i in range(lib().bo_num_scales(self.c)):
            s = Scale()
            lib().bo_get_scale(self.c, i, C.byref(s))
            self.scales.append(s)

    def __del__(self):
        if getattr(self, "c", None):
            lib().bo_destroy(self.c)
            self.c = None

    def set_resize_cpu(self, on=True):
        """Pyramid levels rounded as the reference's resize_cpu option (host pyrDown)."""
        assert lib().bo_set_resize_cpu(self.c, 1 if on else 0) == 0

    def set_image(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        assert img.shape == (self.h, self.w)
        assert lib().bo_set_image(self.c, ptr(img), self.w) == 0
        self._img = img

    def level_image(self, i):
        s = self.scales[i]
        out = np.zeros((s.dataH, s.dataW), np.uint8)
        assert lib().bo_get_level_image(self.c, i, ptr(out)) == 0
        return out

    def filter(self, i):
        s = self.scales[i]
        maps = np.zeros((8, s.fftH, s.fftW), np.float32)
        assert lib().bo_filter_scale(self.c, i, ptr(maps)) == 0
        return maps

    def direct_filter(self, i, nsig=0.0, fp16=False):
        """Precision-form complex-cell maps (FP64 with rounded taps), 8 x fftH x fftW."""
        s = self.scales[i]
        maps = np.zeros((8, s.fftH, s.fftW), np.float32)
        assert lib().bo_direct_filter_scale(self.c, i, nsig, int(fp16), ptr(maps)) == 0
        return maps

    def cells(self, i, maps):
        s = self.scales[i]
        maps = np.ascontiguousarray(maps, np.float32)
        assert maps.shape == (8, s.fftH, s.fftW)
        D = np.zeros((s.dataH, s.dataW), np.float32)
        S = np.zeros((s.dataH, s.dataW), np.float32)
        assert lib().bo_cells_scale(self.c, i, ptr(maps), ptr(D), ptr(S)) == 0
        return D, S

    def nms(self, i, D, cap=8000, count=0):
        D = np.ascontiguousarray(D, np.float32)
        out = np.zeros(cap, KP_DTYPE)
        n = C.c_int(count)
        assert lib().bo_nms_scale(self.c, i, ptr(D), ptr(out), cap, C.byref(n)) == 0
        return out[:min(n.value, cap)], n.value

    def nms_cpu_mode(self, i, D, S, cap=8000):
        D = np.ascontiguousarray(D, np.float32)
        S = np.ascontiguousarray(S, np.float32)
        out = np.zeros(cap, KP_DTYPE)
        n = C.c_int(0)
        assert lib().bo_nms_cpu_mode(self.c, i, ptr(D), ptr(S), ptr(out), cap, C.byref(n)) == 0
        return out[:min(n.value, cap)], n.value

    def detect(self, img, cap=8000, want_maps=False):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros(cap, KP_DTYPE)
        n = C.c_int(0)
        tot = sum(s.dataH * s.dataW for s in self.scales)
        D = np.zeros(tot, np.float32) if want_maps else None
        S = np.zeros(tot, np.float32) if want_maps else None
        assert lib().bo_detect(self.c, ptr(img), self.w, ptr(out), cap, C.byref(n),
                               ptr(D) if want_maps else None,
                               ptr(S) if want_maps else None) == 0
        kps = out[:min(n.value, cap)].copy()
        if not want_maps:
            return kps, n.value
        Ds, Ss, o = [], [], 0
        for s in self.scales:
            A = s.dataH * s.dataW
            Ds.append(D[o:o + A].reshape(s.dataH, s.dataW))
            Ss.append(S[o:o + A].reshape(s.dataH, s.dataW))
            o += A
        return kps, n.value, Ds, Ss


BASE_BANK = np.array([4.0 * 2.0 ** (k / 4.0) for k in range(8)], dtype=np.float64).astype(np.float32)
