// LDS-staged Stockham FFT building blocks for CDNA4 (gfx950), complex64.
//
// A length-N transform (N = 2^k, 8 <= N <= 4096) is computed by N/EPT
// threads that each hold EPT complex values in registers (EPT = 8, 16 or
// 32).  Thread t loads elements x[t + m*N/EPT] (m = 0..EPT-1), runs a
// leading radix-2/4/8 pass followed by radix-8 passes (autosorting Stockham,
// Govindaraju et al. 2008), exchanging through LDS between passes, and ends
// holding X[t + m*N/EPT] in the same register order it loaded from.  That
// symmetry is what lets the row and column kernels fuse pointwise work
// (transmission, propagator, form-factor sum, |psi|^2) between a forward and
// an inverse transform without another trip through LDS or HBM.
//
// Twiddles come from one table w[m] = exp(-2 pi i m / 4096) that the host
// computes in float64 and rounds to float32; a pass of radix R over
// sub-transforms of length Ns reads w[k r 4096 / (Ns R)].
#pragma once

#include <hip/hip_runtime.h>

namespace emsim {

constexpr int kTwN = 4096;   // twiddle table length == largest supported N

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

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

// LDS plane layout of one transform: separate re / im float planes with one
// pad word every 32 so that the power-of-two strided Stockham writes spread
// over the banks.  kStride (odd) is the distance between the planes of
// neighbouring transforms, so that C side-by-side transforms (the column
// kernels) do not fall on the same banks.
template <int N>
struct LdsLayout {
    static constexpr int kWords = N + N / 32;
    static constexpr int kStride = kWords | 1;
    __device__ static __forceinline__ int at(int i) { return i + (i >> 5); }
};

// Radix of the leading pass (the rest are radix 8).
constexpr int lead_radix(int n) {
    return (ilog2(n) % 3 == 1) ? 2 : (ilog2(n) % 3 == 2) ? 4 : 8;
}

// Per-thread copy of the twiddles of every pass after the leading one (those
// are all radix 8).  A worker that runs many transforms with the same t (the
// row and column kernels) loads them once.  fence() makes those loads
// complete where it is called: inside a loop that prefetches its next tile,
// the compiler would otherwise count the twiddle loads as possibly pending
// and, waiting for them, also wait for the prefetch.
template <int N, int EPT>
struct TwCache {
    static constexpr int R0 = lead_radix(N);
    static constexpr int P = (ilog2(N) - ilog2(R0)) / 3;   // radix-8 passes with twiddles
    static constexpr int NB = EPT / 8;                    // butterflies per thread and pass
    float2 w[P > 0 ? P : 1][NB][7];
    static constexpr int pass_of(int ns) { return (ilog2(ns) - ilog2(R0)) / 3; }
    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        int ns = R0;
#pragma unroll
        for (int p = 0; p < P; ++p) {
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const int k = (t + q * (N / EPT)) & (ns - 1);
#pragma unroll
                for (int r = 1; r < 8; ++r) w[p][q][r - 1] = tw[k * r * (kTwN / (ns * 8))];
            }
            ns *= 8;
        }
    }
    __device__ __forceinline__ void fence() const {
#pragma unroll
        for (int p = 0; p < P; ++p)
#pragma unroll
            for (int q = 0; q < NB; ++q)
#pragma unroll
                for (int r = 0; r < 7; ++r) asm volatile("" : : "v"(w[p][q][r].x), "v"(w[p][q][r].y));
    }
};

// Twiddle r of butterfly q (sub-transform index k) of the pass over
// sub-transforms of length Ns: from the table, or from a thread's cache.
template <int N, int EPT, int Ns, int R>
__device__ __forceinline__ float2 tw_get(const float2* __restrict__ tw, int q, int r, int k) {
    return tw[k * r * (kTwN / (Ns * R))];
}
template <int N, int EPT, int Ns, int R>
__device__ __forceinline__ float2 tw_get(const TwCache<N, EPT>& c, int q, int r, int k) {
    static_assert(R == 8, "only the radix-8 passes are cached");
    return c.w[TwCache<N, EPT>::pass_of(Ns)][q][r - 1];
}

// The N twiddles w_N^j of one transform length staged in LDS (p[j], j < N):
// every twiddle of every pass at a low-latency load and no registers.
template <int LEN>
struct TwLds {
    const float2* p;
    // cooperative fill from the table; the caller syncs before use
    __device__ __forceinline__ static void fill(float2* dst, const float2* __restrict__ tw,
                                                int tid, int nthreads) {
        for (int j = tid; j < LEN; j += nthreads) dst[j] = tw[j * (kTwN / LEN)];
    }
};
template <int N, int EPT, int Ns, int R>
__device__ __forceinline__ float2 tw_get(const TwLds<N>& c, int q, int r, int k) {
    return c.p[k * r * (N / (Ns * R))];
}

// One Stockham pass: radix R over sub-transforms of length Ns, EPT values
// per thread.  Before: v[m] = x[t + m N/EPT].  Butterfly j = t + q N/EPT
// (q < EPT/R) reads x[j + r N/R] (= v[q + r EPT/R]) and writes
// y[(j/Ns) Ns R + j%Ns + r Ns].
// SYNC2: the barrier after the reads, which keeps the next pass's writes off
// planes still being read; a caller that alternates two plane sets drops it.
template <int N, int EPT, int R, int Ns, bool INV, class TW, bool SYNC2 = true>
__device__ __forceinline__ void stockham_pass(float2 (&v)[EPT], float* __restrict__ re,
                                              float* __restrict__ im, int t, const TW& tw) {
    constexpr int NB = EPT / R;
    constexpr int S = N / EPT;
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
            for (int r = 1; r < R; ++r) a[r] = mul_tw<INV>(a[r], tw_get<N, EPT, Ns, R>(tw, q, r, k));
        }
        dft<R, INV>(a);
        if constexpr (LAST) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q + r * NB] = a[r];
        } else {
            // at(base + r Ns) == at(base) + at(r Ns): base is a multiple of
            // Ns R plus k < Ns, so adding r Ns never carries across a
            // 32-word pad boundary; the constant part becomes an immediate
            // offset of the LDS instruction
            const int base = L::at((j - k) * R + k);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                re[base + L::at(r * Ns)] = a[r].x;
                im[base + L::at(r * Ns)] = a[r].y;
            }
        }
    }
    if constexpr (!LAST) {
        __syncthreads();
        // at(t + m S) == at(t) + at(m S) for t < S (S a power of two)
        const int base = L::at(t);
#pragma unroll
        for (int m = 0; m < EPT; ++m) v[m] = make_float2(re[base + L::at(m * S)], im[base + L::at(m * S)]);
        if constexpr (SYNC2) __syncthreads();
    }
}

template <int N, int EPT, int Ns, bool INV, class TW>
__device__ __forceinline__ void fft_rest(float2 (&v)[EPT], float* re, float* im, int t,
                                         const TW& tw) {
    if constexpr (Ns < N) {
        stockham_pass<N, EPT, 8, Ns, INV>(v, re, im, t, tw);
        fft_rest<N, EPT, Ns * 8, INV>(v, re, im, t, tw);
    }
}

// LDS exchanges of one transform (every pass but the last writes one)
constexpr int fft_exchanges(int n) { return (ilog2(n) - ilog2(lead_radix(n))) / 3; }

template <int N, int EPT, int Ns, bool INV, int PASS, class TW>
__device__ __forceinline__ void fft_rest_db(float2 (&v)[EPT], float* re0, float* im0, float* re1,
                                            float* im1, int t, const TW& tw) {
    if constexpr (Ns < N) {
        stockham_pass<N, EPT, 8, Ns, INV, TW, false>(v, PASS % 2 ? re1 : re0,
                                                     PASS % 2 ? im1 : im0, t, tw);
        fft_rest_db<N, EPT, Ns * 8, INV, PASS + 1>(v, re0, im0, re1, im1, t, tw);
    }
}

// fft_n with the exchanges alternating between two plane sets, one barrier
// per exchange.  Safe for back-to-back transforms of the same block when
// the exchange count is even: a thread writes a plane set only after the
// barrier of the exchange that followed its last use, which every thread
// passes only once its reads of that set are done.
template <int N, int EPT, bool INV, class TW>
__device__ __forceinline__ void fft_n_db(float2 (&v)[EPT], float* re0, float* im0, float* re1,
                                         float* im1, int t, const TW& tw) {
    static_assert(fft_exchanges(N) % 2 == 0, "alternating planes need an even exchange count");
    constexpr int R0 = lead_radix(N);
    stockham_pass<N, EPT, R0, 1, INV, TW, false>(v, re0, im0, t, tw);
    fft_rest_db<N, EPT, R0, INV, 1>(v, re0, im0, re1, im1, t, tw);
}

// Unnormalised length-N DFT of the values held by the N/EPT threads of one
// row/column group.  Every thread of the block must call it (it contains
// block barriers when N > 8).  tw is the twiddle table (const float2*) or
// this thread's TwCache<N, EPT>.
template <int N, int EPT, bool INV, class TW>
__device__ __forceinline__ void fft_n(float2 (&v)[EPT], float* re, float* im, int t,
                                      const TW& tw) {
    static_assert(N >= 8 && N <= kTwN && (N & (N - 1)) == 0, "unsupported FFT length");
    static_assert(EPT >= 8 && EPT % 8 == 0 && EPT <= N, "EPT must be a multiple of 8");
    constexpr int R0 = lead_radix(N);
    stockham_pass<N, EPT, R0, 1, INV>(v, re, im, t, tw);
    fft_rest<N, EPT, R0, INV>(v, re, im, t, tw);
}

// ---------------------------------------------------------------------------
// sin and cos of a float32 argument.
// |x| <= 8192: Cody-Waite reduction by pi/2 with a three-part float32 split
// of pi/2 (products exact inside FMA), then polynomials on [-pi/4, pi/4]
// (sin r = r + r^3 p(r^2), cos r = 1 - r^2/2 + r^4 q(r^2), p and q of
// degree 2, fitted for least relative error with iterative reweighting
// towards minimax; evaluated in float32 they are within 0.74 ulp (sin) and
// 1.23 ulp (cos) of the exact values over that interval), then quadrant
// selection.  Larger arguments use the ROCm device library's sincosf.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void fast_sincos(float x, float* s_out, float* c_out) {
    if (!(fabsf(x) <= 8192.f)) {
        sincosf(x, s_out, c_out);
        return;
    }
    // pi/2 = P1 + P2 + P3 (P1, P2 float32 with trailing zero bits)
    const float kTwoOverPi = 0.636619772367581343f;
    const float P1 = 1.5703125f;                 // 8 significant bits
    const float P2 = 4.83751296997070312500e-04f;
    const float P3 = 7.54978995489188216e-08f;
    const float kf = rintf(x * kTwoOverPi);
    float r = fmaf(-kf, P1, x);
    r = fmaf(-kf, P2, r);
    r = fmaf(-kf, P3, r);
    const float r2 = r * r;
    float ps = -1.951528392964974e-04f;
    ps = fmaf(ps, r2, 8.332161232829094e-03f);
    ps = fmaf(ps, r2, -1.666665524244309e-01f);
    const float sn = fmaf(ps, r * r2, r);
    float pc = 2.4433154976577498e-05f;
    pc = fmaf(pc, r2, -1.3887316454201937e-03f);
    pc = fmaf(pc, r2, 4.166664555668831e-02f);
    const float cs = fmaf(pc, r2 * r2, fmaf(-0.5f, r2, 1.0f));
    const int q = (int)kf & 3;
    float s = (q & 1) ? cs : sn;
    float c = (q & 1) ? sn : cs;
    if (q == 1 || q == 2) c = -c;
    if (q & 2) s = -s;
    *s_out = s;
    *c_out = c;
}

}  // namespace emsim
