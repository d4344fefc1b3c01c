// fft_device.hpp -- register-resident Stockham FFT for gfx950 wavefronts.
//
// A length-N transform is owned by a group of TPV lanes (TPV = 64 for 512 <= N <= 3584: one
// wavefront per vector; TPV = 128 for 4096: two wavefronts, so that a lane keeps 32 values and
// not 64); every lane keeps E = N/TPV complex values in registers.  Each
// radix-R stage does E/R butterflies per lane entirely in registers; between stages the
// group exchanges data through its own LDS region (N floats: real parts, then imaginary
// parts), so the only cross-lane traffic is one LDS write + read per stage boundary.
//
// Index maps (Stockham autosort, radix list R0..RL, Ns = product of earlier radices):
//   stage input  of butterfly j = lane + TPV*m, leg r : j + r*N/R
//   stage output of butterfly j, leg r              : (j/Ns)*Ns*R + j%Ns + r*Ns
// so the first stage reads element in_index(e) and the last stage produces element
// out_index(e) in register slot e -- both are unit-stride across lanes, which makes global
// loads and stores of a vector coalesced without any extra shuffle.  A group of up to 64
// lanes lies in one wavefront, so its stage boundaries need only wavefront-level
// synchronisation; a 128-lane group synchronises with the workgroup barrier (stage_sync).
//
// Twiddles come from a per-N packed table (TwLayout) computed in FP64 on the host and
// rounded to FP32; the inverse transform uses their conjugates (unnormalised).
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>

namespace bimpdev {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- small DFTs
// out[q] = sum_r in[r] * w^(q*r), w = exp(-+2*pi*i/R), natural output order, in place.
template <bool INV>
__device__ __forceinline__ void dft2(float2* v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <bool INV>
__device__ __forceinline__ void dft3(float2* v) {
    const float kS3 = 0.86602540378443865f;  // sin(pi/3)
    const float2 a = v[0], b = v[1], c = v[2];
    const float2 t1 = cadd(b, c);
    const float2 t2 = make_float2(a.x - 0.5f * t1.x, a.y - 0.5f * t1.y);
    const float2 t3 = cscale(csub(b, c), kS3);
    v[0] = cadd(a, t1);
    v[1] = cadd(t2, mul_mi<INV>(t3));
    v[2] = csub(t2, mul_mi<INV>(t3));
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* v) {
    const float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const float2 t2 = cadd(v[1], v[3]);
    const float2 t3 = mul_mi<INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft5(float2* v) {
    const float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
    const float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
    const float2 a = v[0];
    const float2 t1 = cadd(v[1], v[4]), t3 = csub(v[1], v[4]);
    const float2 t2 = cadd(v[2], v[3]), t4 = csub(v[2], v[3]);
    const float2 A1 = make_float2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const float2 A2 = make_float2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const float2 B1 = make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const float2 B2 = make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    v[0] = cadd(a, cadd(t1, t2));
    v[1] = cadd(A1, mul_mi<INV>(B1));
    v[4] = csub(A1, mul_mi<INV>(B1));
    v[2] = cadd(A2, mul_mi<INV>(B2));
    v[3] = csub(A2, mul_mi<INV>(B2));
}

// X_q = sum_k x_k w^(qk), w = exp(-+2*pi*i/7): with p_k = x_k + x_{7-k} and m_k = x_k - x_{7-k}
// (k = 1..3), X_q = x_0 + sum_k cos(2*pi*q*k/7) p_k -+ i sum_k sin(2*pi*q*k/7) m_k and
// X_{7-q} its partner with the opposite sign of the sine sum.
template <bool INV>
__device__ __forceinline__ void dft7(float2* v) {
    const float c1 = 0.62348980185873353f, c2 = -0.22252093395631440f, c3 = -0.90096886790241913f;
    const float s1 = 0.78183148246802981f, s2 = 0.97492791218182361f, s3 = 0.43388373911755812f;
    const float2 a = v[0];
    const float2 p1 = cadd(v[1], v[6]), m1 = csub(v[1], v[6]);
    const float2 p2 = cadd(v[2], v[5]), m2 = csub(v[2], v[5]);
    const float2 p3 = cadd(v[3], v[4]), m3 = csub(v[3], v[4]);
    // cos(2 pi q k / 7) and sin(2 pi q k / 7) for q, k = 1..3 reduce to c1..c3, +-s1..s3
    const float2 A1 = make_float2(a.x + c1 * p1.x + c2 * p2.x + c3 * p3.x,
                                  a.y + c1 * p1.y + c2 * p2.y + c3 * p3.y);
    const float2 A2 = make_float2(a.x + c2 * p1.x + c3 * p2.x + c1 * p3.x,
                                  a.y + c2 * p1.y + c3 * p2.y + c1 * p3.y);
    const float2 A3 = make_float2(a.x + c3 * p1.x + c1 * p2.x + c2 * p3.x,
                                  a.y + c3 * p1.y + c1 * p2.y + c2 * p3.y);
    const float2 B1 = make_float2(s1 * m1.x + s2 * m2.x + s3 * m3.x, s1 * m1.y + s2 * m2.y + s3 * m3.y);
    const float2 B2 = make_float2(s2 * m1.x - s3 * m2.x - s1 * m3.x, s2 * m1.y - s3 * m2.y - s1 * m3.y);
    const float2 B3 = make_float2(s3 * m1.x - s1 * m2.x + s2 * m3.x, s3 * m1.y - s1 * m2.y + s2 * m3.y);
    v[0] = cadd(a, cadd(p1, cadd(p2, p3)));
    v[1] = cadd(A1, mul_mi<INV>(B1));
    v[6] = csub(A1, mul_mi<INV>(B1));
    v[2] = cadd(A2, mul_mi<INV>(B2));
    v[5] = csub(A2, mul_mi<INV>(B2));
    v[3] = cadd(A3, mul_mi<INV>(B3));
    v[4] = csub(A3, mul_mi<INV>(B3));
}

// w8^q * x for q = 1..3 (w8 = exp(-+i*pi/4))
template <bool INV, int Q>
__device__ __forceinline__ float2 w8(float2 x) {
    const float h = 0.70710678118654752f;
    if (Q == 1) return INV ? make_float2(h * (x.x - x.y), h * (x.x + x.y))
                           : make_float2(h * (x.x + x.y), h * (x.y - x.x));
    if (Q == 2) return mul_mi<INV>(x);
    return INV ? make_float2(-h * (x.x + x.y), h * (x.x - x.y))
               : make_float2(h * (x.y - x.x), -h * (x.x + x.y));
}

template <bool INV>
__device__ __forceinline__ void dft8(float2* v) {
    float2 e[4] = {v[0], v[2], v[4], v[6]};
    float2 o[4] = {v[1], v[3], v[5], v[7]};
    dft4<INV>(e);
    dft4<INV>(o);
    o[1] = w8<INV, 1>(o[1]);
    o[2] = w8<INV, 2>(o[2]);
    o[3] = w8<INV, 3>(o[3]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = cadd(e[q], o[q]);
        v[q + 4] = csub(e[q], o[q]);
    }
}

template <bool INV>
__device__ __forceinline__ void dft16(float2* v) {
    float2 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft8<INV>(e);
    dft8<INV>(o);
    // w16^q, q = 1..7
    const float c[8] = {1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f,
                        0.0f, -0.38268343236508978f, -0.70710678118654752f, -0.92387953251128674f};
    const float s[8] = {0.0f, 0.38268343236508978f, 0.70710678118654752f, 0.92387953251128674f,
                        1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508978f};
#pragma unroll
    for (int q = 1; q < 8; ++q) {
        const float2 w = make_float2(c[q], INV ? s[q] : -s[q]);
        o[q] = cmul(o[q], w);
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        v[q] = cadd(e[q], o[q]);
        v[q + 8] = csub(e[q], o[q]);
    }
}

template <int N1, int N2, bool INV>
__device__ __forceinline__ void dft_pfa(float2* v);

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 24) dft_pfa<8, 3, INV>(v);
    else if constexpr (R == 40) dft_pfa<8, 5, INV>(v);
    else if constexpr (R == 48) dft_pfa<16, 3, INV>(v);
    else if constexpr (R == 56) dft_pfa<8, 7, INV>(v);
    else if constexpr (R == 2) dft2<INV>(v);
    else if constexpr (R == 3) dft3<INV>(v);
    else if constexpr (R == 4) dft4<INV>(v);
    else if constexpr (R == 5) dft5<INV>(v);
    else if constexpr (R == 7) dft7<INV>(v);
    else if constexpr (R == 8) dft8<INV>(v);
    else if constexpr (R == 16) dft16<INV>(v);
    else static_assert(R == 2, "unsupported radix");
}

// q with a*q = 1 (mod m), a and m coprime
__host__ __device__ constexpr int inv_mod(int a, int m) {
    for (int q = 1; q < m; ++q)
        if (a * q % m == 1) return q;
    return 1;
}

// Prime-factor (Good-Thomas) DFT of length N = N1*N2, gcd(N1, N2) = 1: with input index
// n = (N2 n1 + N1 n2) mod N and output index k = (N2 q1 k1 + N1 q2 k2) mod N, where
// q1 = N2^-1 mod N1 and q2 = N1^-1 mod N2, w_N^(nk) = w_N1^(n1 k1) w_N2^(n2 k2), so N2
// length-N1 DFTs followed by N1 length-N2 DFTs give the transform with no twiddles between
// them.  All indices are compile-time: the permutations are register renamings.
template <int N1, int N2, bool INV>
__device__ __forceinline__ void dft_pfa(float2* v) {
    constexpr int N = N1 * N2;
    constexpr int q1 = inv_mod(N2 % N1, N1), q2 = inv_mod(N1 % N2, N2);
    static_assert(N2 * q1 % N1 == 1 && N1 * q2 % N2 == 1, "prime-factor lengths are coprime");
    float2 y[N];  // y[k1 * N2 + n2]
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
        float2 a[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) a[n1] = v[(N2 * n1 + N1 * n2) % N];
        dft<N1, INV>(a);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) y[k1 * N2 + n2] = a[k1];
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        float2 b[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) b[n2] = y[k1 * N2 + n2];
        dft<N2, INV>(b);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) v[(N2 * q1 * k1 + N1 * q2 * k2) % N] = b[k2];
    }
}

// ---------------------------------------------------------------- plans
template <int... Rs>
struct RadList {};

template <int N>
struct FFTCfg;

#define BIMP_FFT_CFG(N_, TPV_, ...)                               \
    template <>                                                    \
    struct FFTCfg<N_> {                                            \
        static constexpr int N = N_;                               \
        static constexpr int TPV = TPV_;                           \
        static constexpr int E = N_ / TPV_;                        \
        using Rad = RadList<__VA_ARGS__>;                          \
    };
BIMP_FFT_CFG(16, 4, 4, 4)
BIMP_FFT_CFG(32, 8, 4, 4, 2)
BIMP_FFT_CFG(64, 16, 4, 4, 4)
BIMP_FFT_CFG(128, 16, 8, 8, 2)
BIMP_FFT_CFG(256, 32, 8, 8, 4)
BIMP_FFT_CFG(512, 64, 8, 8, 8)
BIMP_FFT_CFG(1024, 64, 16, 16, 4)
// The lengths with an odd factor take it in a prime-factor first stage (24 = 8*3, 40 = 8*5,
// 48 = 16*3, 56 = 8*7): three stages and two LDS exchanges instead of four and three.
BIMP_FFT_CFG(1536, 64, 24, 8, 8)
BIMP_FFT_CFG(2048, 64, 16, 16, 8)
BIMP_FFT_CFG(2560, 64, 40, 8, 8)
BIMP_FFT_CFG(3072, 64, 48, 8, 8)
BIMP_FFT_CFG(3584, 64, 56, 8, 8)
BIMP_FFT_CFG(4096, 128, 16, 16, 16)
#undef BIMP_FFT_CFG

template <class L>
struct RadInfo;
template <int R0, int... Rs>
struct RadInfo<RadList<R0, Rs...>> {
    static constexpr int first = R0;
    static constexpr int last = RadInfo<RadList<Rs...>>::last;
};
template <int R0>
struct RadInfo<RadList<R0>> {
    static constexpr int first = R0;
    static constexpr int last = R0;
};

// Packed per-stage twiddle table.  Stage t (radix R, Ns = product of the earlier radices)
// keeps, for every k in [0, Ns), the R-1 twiddles exp(-2*pi*i*r*k/(Ns*R)), r = 1..R-1, in a
// row padded to an even length, so a butterfly fetches its twiddles as float4 pairs from one
// base address.  Stage 0 (Ns = 1) needs none.  Stages are concatenated in plan order.
__host__ __device__ constexpr int tw_row(int R) { return (R - 1 + 1) / 2 * 2; }

template <class L, int Ns = 1, bool First = true>
struct TwLayout;
template <int R, int... Rest, int Ns, bool First>
struct TwLayout<RadList<R, Rest...>, Ns, First> {
    static constexpr int size = (First ? 0 : Ns * tw_row(R)) + TwLayout<RadList<Rest...>, Ns * R, false>::size;
};
template <int Ns, bool First>
struct TwLayout<RadList<>, Ns, First> {
    static constexpr int size = 0;
};

// element index held in register slot e before the first stage / after the last stage
template <int N>
__device__ __forceinline__ int in_index(int e, int lane) {
    using C = FFTCfg<N>;
    constexpr int R = RadInfo<typename C::Rad>::first;
    return lane + C::TPV * (e / R) + (e % R) * (N / R);
}
template <int N>
__device__ __forceinline__ int out_index(int e, int lane) {
    using C = FFTCfg<N>;
    constexpr int R = RadInfo<typename C::Rad>::last;
    return lane + C::TPV * (e / R) + (e % R) * (N / R);
}

// LDS ordering between the stages of one group of TPV lanes.  A group of up to 64 lanes lies
// in one wavefront (TPV divides 64) and owns its LDS region, so a wavefront-scope
// release/acquire around a wave barrier orders the exchange (the same sequence HIP uses for
// __syncwarp).  A 128-lane group spans two wavefronts and takes the workgroup barrier: every
// wavefront of a workgroup that runs such transforms runs the same number of them.
template <int TPV>
__device__ __forceinline__ void stage_sync() {
    if constexpr (TPV > 64) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Twiddles copied into LDS by a workgroup that runs many transforms (tw_to_lds below): each
// stage's Ns rows of P/2 float4 are stored transposed, float4 q of row k at q*Ns + k, so the
// lanes of a wavefront (consecutive k) read consecutive float4s.
struct LdsTw {
    const __attribute__((address_space(3))) float4* p;
};
// Twiddles read from global memory (L1/L2 hits): a global-address-space pointer, so that the
// loads are global_* (vmcnt only) and not flat_*, which would also hold up the LDS waits of
// the exchanges.
typedef const __attribute__((address_space(1))) float2* GlobalTw;

template <int N, bool INV, int Ns, int R, class TW>
__device__ __forceinline__ void butterflies(float2 (&v)[FFTCfg<N>::E], int lane, TW tw) {
    using C = FFTCfg<N>;
    constexpr int P = tw_row(R);
#pragma unroll
    for (int m = 0; m < C::E / R; ++m) {
        if constexpr (Ns > 1) {
            const int j = lane + C::TPV * m;
            const int k = j % Ns;
            float4 t[P / 2];
            if constexpr (std::is_same<TW, LdsTw>::value) {
#pragma unroll
                for (int q = 0; q < P / 2; ++q)
                    t[q] = *(const float4*)(tw.p + q * Ns + k);  // generic: address space inferred
            } else {
                typedef float f4v __attribute__((ext_vector_type(4)));
                const __attribute__((address_space(1))) f4v* row =
                    (const __attribute__((address_space(1))) f4v*)(tw + k * P);
#pragma unroll
                for (int q = 0; q < P / 2; ++q) {
                    const f4v x = row[q];
                    t[q] = make_float4(x.x, x.y, x.z, x.w);
                }
            }
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const float4 tq = t[(r - 1) / 2];
                float2 w = ((r - 1) & 1) ? make_float2(tq.z, tq.w) : make_float2(tq.x, tq.y);
                if (INV) w.y = -w.y;
                v[m * R + r] = cmul(v[m * R + r], w);
            }
        }
        dft<R, INV>(&v[m * R]);
    }
}

// LDS floats one group's exchange region needs, and the address of element i in it.  After
// the first stage lane j holds its R outputs j*R .. j*R+R-1.  When R <= VMAX (a parameter of
// the transform) they leave as R/4 float4 writes: consecutive lanes' 16-byte writes 4R bytes
// apart, one instruction for four scalar writes and no per-element address arithmetic.
// The prime-factor radices (R > 16) always take the float4 form.
// Otherwise each is a scalar write whose address XOR-swizzles the bank bits of its 32-word row
// with the row index (a bijection within the row: it spreads the stride-R writes over all
// banks and keeps the reads of 32 consecutive elements conflict-free, at no extra LDS).  The
// float4 form is faster for R = 16 in the column and forward row passes and slower in the
// inverse row passes, which therefore keep the swizzle for R = 16.  Later exchanges write runs
// of >= R consecutive lanes and keep the plain layout.
template <int Ns, int R, int VMAX>
inline constexpr bool kVecFirst = Ns == 1 && R % 4 == 0 && (R <= VMAX || R > 16);

template <int N>
__host__ __device__ constexpr int fft_lds_floats() { return N; }
template <int Ns, int R, int VMAX>
__device__ __forceinline__ int lds_pad(int i) {
    if constexpr (Ns == 1 && !kVecFirst<Ns, R, VMAX>)
        return (i & ~31) | ((i ^ (i >> 5)) & 31);
    else
        return i;
}

// The exchange moves real parts, then imaginary parts, through an LDS region of
// fft_lds_floats<N>() floats: a register slot keeps its old imaginary part until the second
// half has been written, so half the LDS of a complex exchange suffices.
template <int N, int Ns, int R, int R2, int VMAX>
__device__ __forceinline__ void exchange(float2 (&v)[FFTCfg<N>::E], float* lds, int lane) {
    using C = FFTCfg<N>;
    if constexpr (kVecFirst<Ns, R, VMAX>) {
#pragma unroll
        for (int m = 0; m < C::E / R; ++m) {
            float4* dst = reinterpret_cast<float4*>(lds + (lane + C::TPV * m) * R);
#pragma unroll
            for (int q = 0; q < R / 4; ++q)
                dst[q] = make_float4(v[m * R + 4 * q].x, v[m * R + 4 * q + 1].x,
                                     v[m * R + 4 * q + 2].x, v[m * R + 4 * q + 3].x);
        }
    } else {
#pragma unroll
        for (int m = 0; m < C::E / R; ++m) {
            const int j = lane + C::TPV * m;
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lds_pad<Ns, R, VMAX>(base + r * Ns)] = v[m * R + r].x;
        }
    }
    stage_sync<C::TPV>();
    float im[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) im[e] = v[e].y;
#pragma unroll
    for (int m = 0; m < C::E / R2; ++m) {
        const int j = lane + C::TPV * m;
#pragma unroll
        for (int r = 0; r < R2; ++r) v[m * R2 + r].x = lds[lds_pad<Ns, R, VMAX>(j + r * (N / R2))];
    }
    stage_sync<C::TPV>();
    if constexpr (kVecFirst<Ns, R, VMAX>) {
#pragma unroll
        for (int m = 0; m < C::E / R; ++m) {
            float4* dst = reinterpret_cast<float4*>(lds + (lane + C::TPV * m) * R);
#pragma unroll
            for (int q = 0; q < R / 4; ++q)
                dst[q] = make_float4(im[m * R + 4 * q], im[m * R + 4 * q + 1],
                                     im[m * R + 4 * q + 2], im[m * R + 4 * q + 3]);
        }
    } else {
#pragma unroll
        for (int m = 0; m < C::E / R; ++m) {
            const int j = lane + C::TPV * m;
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lds_pad<Ns, R, VMAX>(base + r * Ns)] = im[m * R + r];
        }
    }
    stage_sync<C::TPV>();
#pragma unroll
    for (int m = 0; m < C::E / R2; ++m) {
        const int j = lane + C::TPV * m;
#pragma unroll
        for (int r = 0; r < R2; ++r) v[m * R2 + r].y = lds[lds_pad<Ns, R, VMAX>(j + r * (N / R2))];
    }
    stage_sync<C::TPV>();
}

template <int N, bool INV, int Ns, class L, int VMAX>
struct Stages;
template <int N, bool INV, int Ns, int R, int... Rest, int VMAX>
struct Stages<N, INV, Ns, RadList<R, Rest...>, VMAX> {
    // tw points at this stage's rows (unused for the first stage)
    template <class TW>
    static __device__ __forceinline__ void run(float2 (&v)[FFTCfg<N>::E], float* lds, int lane,
                                               TW tw) {
        butterflies<N, INV, Ns, R>(v, lane, tw);
        if constexpr (sizeof...(Rest) > 0) {
            exchange<N, Ns, R, RadInfo<RadList<Rest...>>::first, VMAX>(v, lds, lane);
            constexpr int next_ns = Ns * R;
            constexpr int adv = Ns > 1 ? Ns * tw_row(R) : 0;  // float2
            if constexpr (std::is_same<TW, LdsTw>::value)
                Stages<N, INV, next_ns, RadList<Rest...>, VMAX>::run(v, lds, lane, LdsTw{tw.p + adv / 2});
            else
                Stages<N, INV, next_ns, RadList<Rest...>, VMAX>::run(v, lds, lane, tw + adv);
        }
    }
};

// Copies the packed table of FFTCfg<N> (global, row-major) into LDS in the transposed layout
// LdsTw reads; tw_lds_bytes<N>() bytes.  Threads tid = 0..nthr-1 of the workgroup share it.
template <class L, int Ns = 1, bool First = true>
struct TwCopy;
template <int R, int... Rest, int Ns, bool First>
struct TwCopy<RadList<R, Rest...>, Ns, First> {
    static __device__ __forceinline__ void run(const float4* __restrict__ src, float4* dst,
                                               int tid, int nthr) {
        if constexpr (!First) {
            constexpr int P4 = tw_row(R) / 2;
            for (int i = tid; i < Ns * P4; i += nthr) {
                const int k = i / P4, q = i - k * P4;
                dst[q * Ns + k] = src[i];
            }
            TwCopy<RadList<Rest...>, Ns * R, false>::run(src + Ns * P4, dst + Ns * P4, tid, nthr);
        } else {
            TwCopy<RadList<Rest...>, Ns * R, false>::run(src, dst, tid, nthr);
        }
    }
};
template <int Ns, bool First>
struct TwCopy<RadList<>, Ns, First> {
    static __device__ __forceinline__ void run(const float4*, float4*, int, int) {}
};

template <int N>
__host__ __device__ constexpr size_t tw_lds_bytes() {
    return (size_t)TwLayout<typename FFTCfg<N>::Rad>::size * sizeof(float2);
}
template <int N>
__device__ __forceinline__ void tw_to_lds(const float2* tw, float4* dst, int tid, int nthr) {
    TwCopy<typename FFTCfg<N>::Rad>::run(reinterpret_cast<const float4*>(tw), dst, tid, nthr);
}

// The twiddle pointer is passed through an empty asm statement so that a kernel running
// several transforms in a loop reloads twiddles (L1 hits) instead of keeping every one of
// them live in registers across iterations.
template <int N, bool INV, int VMAX = 16>
__device__ __forceinline__ void fft(float2 (&v)[FFTCfg<N>::E], float* lds, int lane,
                                    const float2* tw) {
    GlobalTw g = (GlobalTw)tw;
    asm volatile("" : "+s"(g));
    Stages<N, INV, 1, typename FFTCfg<N>::Rad, VMAX>::run(v, lds, lane, g);
}
template <int N, bool INV, int VMAX = 16>
__device__ __forceinline__ void fft(float2 (&v)[FFTCfg<N>::E], float* lds, int lane, LdsTw tw) {
    asm volatile("" : "+s"(tw.p));
    Stages<N, INV, 1, typename FFTCfg<N>::Rad, VMAX>::run(v, lds, lane, tw);
}

}  // namespace bimpdev
