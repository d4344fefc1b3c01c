"""FP64 numpy restatement of the differentiation form: forward and reference gradient.

Test infrastructure (the checker of bimp_diff_* and the libtorch autograd Function).  The
forward restates, in FP64 on float images: the float pyramid (5x5 binomial / 256, reflect-101,
no rounding), clamp padding (convolutionFFT2D.cuh:56-85), the Gabor bank convolution over the
FH x FW torus (cuda_bimp.cu:304-331, :486-507), complex-cell magnitudes, and the D map of the
end-stop / inhibition stencil with the reference's linear-index reads (cuda_bimp.cuh:436-517).
The gradient is an independent adjoint: stencil contributions are scattered with np.add.at
(the library gathers), the convolution adjoints use numpy's FFT.  It is itself pinned by FP64
central differences in tests/test_diff_ref.py.
"""
import numpy as np

import oracle_lib as O

KNUM = 8


def gabor(s, o, ratio, gamma):
    """Ke, Ko (with Ke's DC offset removed) and their derivatives by (ratio, gamma), FP64."""
    lam = float(s.lam)
    sigma = lam * ratio
    th = o * np.pi / 8
    kW, kX = s.kW, s.kX
    y, x = np.mgrid[0:kW, 0:kW].astype(np.float64) - kX
    xr = x * np.cos(th) + y * np.sin(th)
    yr = y * np.cos(th) - x * np.sin(th)
    q = xr * xr + gamma * yr * yr
    env = np.exp(-q / (2 * sigma * sigma))
    f = 2 * np.pi * xr / lam
    fs2 = float(s.fs) ** 2
    e = env * np.cos(f)
    ko = env * np.cos(f - np.pi / 2)
    ke = e - e.sum() / fs2
    denv_r = env * q * lam / sigma ** 3
    denv_g = -env * yr * yr / (2 * sigma * sigma)
    dke_r = denv_r * np.cos(f)
    dke_r -= dke_r.sum() / fs2
    dke_g = denv_g * np.cos(f)
    dke_g -= dke_g.sum() / fs2
    return ke, ko, (dke_r, denv_r * np.cos(f - np.pi / 2)), (dke_g, denv_g * np.cos(f - np.pi / 2))


def _pyr_matrix(n):
    """Dense 1D pyramid operator ((n+1)//2 x n), reflect-101."""
    dn = (n + 1) // 2
    k = np.array([1, 4, 6, 4, 1], np.float64) / 16
    A = np.zeros((dn, n))
    for d in range(dn):
        for i in range(-2, 3):
            t = 2 * d + i
            if t < 0:
                t = -t
            if t >= n:
                t = 2 * (n - 1) - t
            A[d, t] += k[i + 2]
    return A


def _pad_index(n, F, k):
    i = np.arange(F)
    return np.where(i < n, i, np.where(i < n + k, n - 1, 0))


def _half_away(v):
    return int(np.sign(v) * np.floor(abs(v) + 0.5))


def _offsets(lam):
    """(end-stop a, b, inhibition a, b, radial a, b) per orientation (oracle offsets; radial
    halves rounded half away from zero)."""
    out = []
    for o in range(KNUM):
        ea, eb = O.end_stop_offsets(lam, o)
        ia, ib = O.inhibition_offsets(lam, o)
        out.append((ea, eb, ia, ib, _half_away(0.5 * ia), _half_away(0.5 * ib)))
    return out


class DiffRef:
    def __init__(self, w, h, lambdas, input_scale=255.0 / 65536.0):
        self.w, self.h = w, h
        self.scales = [O.scale_params(float(l), w, h) for l in sorted(lambdas)]
        self.sc = 255.0 * np.float64(np.float32(input_scale))
        self.maxlevel = max(s.level for s in self.scales)

    def _levels(self, img):
        lev = [np.asarray(img, np.float64)]
        mats = []
        for _ in range(self.maxlevel):
            L = lev[-1]
            Ay, Ax = _pyr_matrix(L.shape[0]), _pyr_matrix(L.shape[1])
            mats.append((Ay, Ax))
            lev.append(Ay @ L @ Ax.T)
        return lev, mats

    def _scale(self, s, L, ratio, gamma):
        FH, FW, P = s.fftH, s.fftW, s.fftH * s.fftW
        cy, cx = _pad_index(s.dataH, FH, s.kX), _pad_index(s.dataW, FW, s.kX)
        Pd = self.sc * L[np.ix_(cy, cx)]
        Dh = np.fft.fft2(Pd)
        Q, Z, taps = [], [], []
        for o in range(KNUM):
            ke, ko, dr, dg = gabor(s, o, ratio, gamma)
            pad = np.zeros((FH, FW), complex)
            for r in range(s.kW):
                for c in range(s.kW):
                    pad[(r - s.kX) % FH, (c - s.kX) % FW] += ke[r, c] + 1j * ko[r, c]
            q = np.fft.fft2(pad)
            Q.append(q)
            Z.append(np.fft.ifft2(Dh * q))
            taps.append((dr, dg))
        C = np.abs(np.stack(Z))
        return dict(Pd=Pd, Dh=Dh, Q=Q, Z=Z, C=C.reshape(KNUM, P), taps=taps, cy=cy, cx=cx)

    def _stencil(self, s, Cf):
        """D map and the per-term data needed by the adjoint (flat linear indices)."""
        H, W, FW, P = s.dataH, s.dataW, s.fftW, s.fftH * s.fftW
        y, x = np.mgrid[0:H, 0:W]
        g = (y * FW + x).ravel()
        yy, xx = y.ravel(), x.ravel()
        D = np.zeros(H * W)
        terms = []
        for o, (ea, eb, ia, ib, ra, rb) in enumerate(_offsets(s.lam)):
            m = Cf[o]
            m2 = Cf[(o + 4) % KNUM]
            ve = (2 * abs(eb) <= xx) & (xx <= W - 2 * abs(eb)) & (2 * abs(ea) <= yy) & (yy <= H - 2 * abs(ea))
            d1 = -2 * ea * FW + 2 * eb
            i0, i1, i2 = g, (g + d1) % P, (g - d1) % P
            v = m[i0] - 0.5 * m[i1] - 0.5 * m[i2]
            mds = ve & (v > 0)
            D += np.where(mds, v, 0)
            vi = (2 * abs(ib) <= xx) & (xx <= W - 2 * abs(ib)) & (2 * abs(ia) <= yy) & (yy <= H - 2 * abs(ia))
            t = -ia * FW - ib
            it = (g + t) % P
            vt = m[it] - m[i0]
            mt = vi & (vt > 0)
            rr = ra * FW + rb
            ip, im = (g + rr) % P, (g - rr) % P
            v1 = m[i0] - 16 * m2[ip]
            m1 = v1 > 0
            v2 = np.where(m1, v1, 0) + m[i0] - 16 * m2[im]
            mr = vi & (v2 > 0)
            D -= np.where(mt, vt, 0) + np.where(mr, v2, 0)
            terms.append(dict(o=o, i0=i0, i1=i1, i2=i2, it=it, ip=ip, im=im, mds=mds, mt=mt,
                              m1=m1 & mr, mr=mr))
        return D.reshape(H, W), terms

    def forward(self, img, ratio=0.56, gamma=0.5):
        lev, _ = self._levels(img)
        out = []
        for s in self.scales:
            f = self._scale(s, lev[s.level], ratio, gamma)
            out.append(self._stencil(s, f["C"])[0])
        return out

    def loss(self, img, targets, ratio=0.56, gamma=0.5):
        return sum(float(((d - t) ** 2).sum()) for d, t in zip(self.forward(img, ratio, gamma), targets))

    def gradient(self, img, targets, ratio=0.56, gamma=0.5):
        """(loss, d loss / d image, d loss / d (ratio, gamma)) for loss = sum_s |D_s - T_s|^2."""
        lev, mats = self._levels(img)
        glev = [np.zeros_like(L) for L in lev]
        gth = np.zeros(2)
        loss = 0.0
        for s, T in zip(self.scales, targets):
            f = self._scale(s, lev[s.level], ratio, gamma)
            D, terms = self._stencil(s, f["C"])
            loss += float(((D - T) ** 2).sum())
            gD = (2 * (D - T)).ravel()
            P = s.fftH * s.fftW
            gC = np.zeros((KNUM, P))
            for tm in terms:
                o, o2 = tm["o"], (tm["o"] + 4) % KNUM
                a = np.where(tm["mds"], gD, 0)
                np.add.at(gC[o], tm["i0"], a)
                np.add.at(gC[o], tm["i1"], -0.5 * a)
                np.add.at(gC[o], tm["i2"], -0.5 * a)
                b = np.where(tm["mt"], gD, 0)
                np.add.at(gC[o], tm["it"], -b)
                np.add.at(gC[o], tm["i0"], b)
                r2 = np.where(tm["mr"], gD, 0)
                r1 = np.where(tm["m1"], gD, 0)
                np.add.at(gC[o], tm["i0"], -(r2 + r1))
                np.add.at(gC[o2], tm["ip"], 16 * r1)
                np.add.at(gC[o2], tm["im"], 16 * r2)
            FH, FW = s.fftH, s.fftW
            acc = np.zeros((FH, FW), complex)
            for o in range(KNUM):
                Z = f["Z"][o]
                C = np.abs(Z)
                gz = np.where(C > 0, gC[o].reshape(FH, FW) / np.where(C > 0, C, 1), 0) * Z
                G = np.fft.fft2(gz)
                acc += np.conj(f["Q"][o]) * G
                Y = np.fft.ifft2(G * np.conj(f["Dh"]))
                dr, dg = f["taps"][o]
                ry = (np.arange(s.kW) - s.kX) % FH
                rx = (np.arange(s.kW) - s.kX) % FW
                Yt = Y[np.ix_(ry, rx)]
                gth[0] += float((Yt.real * dr[0] + Yt.imag * dr[1]).sum())
                gth[1] += float((Yt.real * dg[0] + Yt.imag * dg[1]).sum())
            gPd = np.real(np.fft.ifft2(acc)) * self.sc
            gL = np.zeros((s.dataH, s.dataW))
            np.add.at(gL, (f["cy"][:, None], f["cx"][None, :]), gPd)
            glev[s.level][: s.dataH, : s.dataW] += gL  # the data is the level's top-left crop
        for l in range(self.maxlevel, 0, -1):
            Ay, Ax = mats[l - 1]
            glev[l - 1] += Ay.T @ glev[l] @ Ax
        return loss, glev[0], gth
