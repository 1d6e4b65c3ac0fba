// LDS-staged Stockham FFT building blocks for CDNA4 (gfx950), complex64.
//
// A length-N transform (N = 2^k, 8 <= N <= 4096) is computed by N/8 threads
// that each hold 8 complex values in registers.  Thread t loads elements
// x[t + m*N/8] (m = 0..7), runs a leading radix-2/4/8 pass followed by
// radix-8 passes (autosorting Stockham, Govindaraju et al. 2008), exchanging
// through LDS between passes, and ends holding X[t + m*N/8] in the same
// register order it loaded from.  That symmetry is what lets the row and
// column kernels fuse pointwise work (transmission, propagator, form-factor
// sum, |psi|^2) between a forward and an inverse transform without another
// trip through LDS or HBM.
//
// Twiddles come from one table w[m] = exp(-2 pi i m / 4096) that the host
// computes in float64 and rounds to float32; a pass of radix R over
// sub-transforms of length Ns reads w[k r 4096 / (Ns R)].
#pragma once

#include <hip/hip_runtime.h>

namespace emsim {

constexpr int kTwN = 4096;   // twiddle table length == largest supported N
constexpr int kEPT = 8;      // complex elements per thread

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a * w (forward) or a * conj(w) (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_tw(float2 a, float2 w) {
    if constexpr (!INV) return make_float2(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
    else return make_float2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}

// a * (-i) forward, a * (+i) inverse
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    if constexpr (!INV) return make_float2(a.y, -a.x);
    else return make_float2(-a.y, a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2* a) {
    const float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    const float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2* a) { dft4<INV>(a[0], a[1], a[2], a[3]); }

template <bool INV>
__device__ __forceinline__ void dft8(float2* a) {
    float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
    float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float c = 0.707106781186547524f;
    if constexpr (!INV) {
        o1 = make_float2(c * (o1.x + o1.y), c * (o1.y - o1.x));     // * e^{-i pi/4}
        o3 = make_float2(c * (o3.y - o3.x), -c * (o3.x + o3.y));    // * e^{-3i pi/4}
    } else {
        o1 = make_float2(c * (o1.x - o1.y), c * (o1.x + o1.y));     // * e^{+i pi/4}
        o3 = make_float2(-c * (o3.x + o3.y), c * (o3.x - o3.y));    // * e^{+3i pi/4}
    }
    o2 = mul_mi<INV>(o2);
    a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
    a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
    a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
    a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
    if constexpr (R == 2) dft2<INV>(a);
    else if constexpr (R == 4) dft4<INV>(a);
    else dft8<INV>(a);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// LDS row layout.  Elements are stored as two float planes (re, im) with one
// pad word every 32 so that the power-of-two strided Stockham writes spread
// over the LDS banks.
template <int N>
struct LdsLayout {
    static constexpr int kWords = N + N / 32;
    __device__ static __forceinline__ int at(int i) { return i + (i >> 5); }
};

// One Stockham pass: radix R over sub-transforms of length Ns.
// Before: v[m] = x[t + m N/8].  Butterfly j = t + q N/8 reads x[j + r N/R]
// (= v[q + r 8/R]) and writes y[(j/Ns) Ns R + j%Ns + r Ns].
template <int N, int R, int Ns, bool INV>
__device__ __forceinline__ void stockham_pass(float2 (&v)[kEPT], float* __restrict__ re,
                                              float* __restrict__ im, int t,
                                              const float2* __restrict__ tw) {
    constexpr int NB = kEPT / R;
    constexpr int S = N / kEPT;
    constexpr bool LAST = (Ns * R == N);
    using L = LdsLayout<N>;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int j = t + q * S;
        const int k = j & (Ns - 1);
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[q + r * NB];
        if constexpr (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) a[r] = mul_tw<INV>(a[r], tw[k * r * (kTwN / (Ns * R))]);
        }
        dft<R, INV>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q + r * NB] = a[r];
        } else {
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int idx = L::at(base + r * Ns);
                re[idx] = a[r].x;
                im[idx] = a[r].y;
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
#pragma unroll
        for (int m = 0; m < kEPT; ++m) {
            const int idx = L::at(t + m * S);
            v[m] = make_float2(re[idx], im[idx]);
        }
        __syncthreads();
    }
}

template <int N, int Ns, bool INV>
__device__ __forceinline__ void fft_rest(float2 (&v)[kEPT], float* re, float* im, int t,
                                         const float2* __restrict__ tw) {
    if constexpr (Ns < N) {
        stockham_pass<N, 8, Ns, INV>(v, re, im, t, tw);
        fft_rest<N, Ns * 8, INV>(v, re, im, t, tw);
    }
}

// Unnormalised length-N DFT of the values held by the N/8 threads of one
// row/column group.  Every thread of the block must call it (it contains
// block barriers when N > 8).
template <int N, bool INV>
__device__ __forceinline__ void fft_n(float2 (&v)[kEPT], float* re, float* im, int t,
                                      const float2* __restrict__ tw) {
    static_assert(N >= 8 && N <= kTwN && (N & (N - 1)) == 0, "unsupported FFT length");
    constexpr int L = ilog2(N);
    constexpr int R0 = (L % 3 == 1) ? 2 : (L % 3 == 2) ? 4 : 8;
    stockham_pass<N, R0, 1, INV>(v, re, im, t, tw);
    fft_rest<N, R0, INV>(v, re, im, t, tw);
}

}  // namespace emsim
