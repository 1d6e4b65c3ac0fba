// DFTs of any length n (8 <= n <= 2048) by Bluestein's chirp-z algorithm on
// the power-of-two FFT kernels, for the reference-shaped API calls at lateral
// sizes that are not powers of two (the reference's FFTW plans take any n,
// emsim/backend/fftw_ext/src/WavePropagator.cpp:25-33, SliceBuilder.cpp:41-55).
//
// With w = exp(-2 pi i / n), j k = (j^2 + k^2 - (k - j)^2) / 2 gives
//   X_k = sum_j x_j w^(j k) = c_k sum_j (x_j c_j) conj(c_(k-j)),
//   c_m = exp(-i pi m^2 / n),
// a convolution, done circularly at a power-of-two length M >= 2n - 1:
//   X = c . IFFT_M(FFT_M(x . c, zero-padded) . B) / M,
// B = FFT_M(b), b_m = conj(c_m) at m and at M - m (0 <= m < n).  The inverse
// transform uses conj(c) in place of c.  c and B are computed on the host in
// float64 (m^2 reduced modulo 2n first) and rounded once.
//
// A batch of "lines" (rows or columns of a batch of 2-D arrays) is
// transformed at a time: line r's element k lives at LineMap::at(r, k).
// The line buffers are (lines, M) complex64, contiguous, so the power-of-two
// row FFT kernels transform them in place.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"
#include "kernels_wave.hpp"

namespace emsim {

struct LineMap {
    int64_t per_img;   // lines per image
    int64_t img;       // image stride (elements)
    int64_t line;      // stride between lines of one image
    int64_t elem;      // stride between a line's elements
    __device__ __forceinline__ int64_t at(int64_t r, int64_t k) const {
        const int64_t q = r / per_img;
        return q * img + (r - q * per_img) * line + k * elem;
    }
};

// How k_chirp_in reads a line:
//   kComplex: n complex values;
//   kReal:    n real values (imaginary part zero);
//   kHalf:    the half spectrum (n/2 + 1 complex values) of a real line,
//             extended to n values by Hermitian symmetry the way a C2R
//             transform reads it (imaginary parts of bin 0 and, for even
//             n, bin n/2 ignored).
enum ChirpIn { kComplex = 0, kReal = 1, kHalf = 2 };

// work[r][j] = (line r)[j] * chirp[j] for j < n, 0 for n <= j < M.
template <int MODE>
__global__ __launch_bounds__(256) void k_chirp_in(const void* __restrict__ in, LineMap lm,
                                                  int64_t lines, int n, int M,
                                                  const float2* __restrict__ chirp,
                                                  float2* __restrict__ work) {
    const int64_t total = lines * (int64_t)M;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / M;
        const int j = (int)(t - r * M);
        float2 v = make_float2(0.f, 0.f);
        if (j < n) {
            float2 x;
            if constexpr (MODE == kReal) {
                x = make_float2(((const float*)in)[lm.at(r, j)], 0.f);
            } else if constexpr (MODE == kHalf) {
                const float2* src = (const float2*)in;
                if (2 * j <= n) {
                    x = src[lm.at(r, j)];
                    if (j == 0 || 2 * j == n) x.y = 0.f;
                } else {
                    const float2 y = src[lm.at(r, n - j)];
                    x = make_float2(y.x, -y.y);
                }
            } else {
                x = ((const float2*)in)[lm.at(r, j)];
            }
            v = cmul(x, chirp[j]);
        }
        work[t] = v;
    }
}

// work[r][j] *= table[j]
__global__ __launch_bounds__(256) void k_line_mul(float2* __restrict__ work, int64_t lines, int M,
                                                  const float2* __restrict__ table) {
    const int64_t total = lines * (int64_t)M;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int j = (int)(t % M);
        work[t] = cmul(work[t], table[j]);
    }
}

// (line r)[k] = work[r][k] * chirp[k] * scale for k < n_out (<= n).
// REAL_OUT: the real part only, clipped below at clip_lo (a C2R transform
// followed by the potential's clip).
template <bool REAL_OUT>
__global__ __launch_bounds__(256) void k_chirp_out(const float2* __restrict__ work,
                                                   int64_t lines, int n_out, int M,
                                                   const float2* __restrict__ chirp, float scale,
                                                   float clip_lo, void* __restrict__ out,
                                                   LineMap lm) {
    const int64_t total = lines * (int64_t)n_out;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int64_t r = t / n_out;
        const int k = (int)(t - r * n_out);
        const float2 v = cmul(work[r * M + k], chirp[k]);
        if constexpr (REAL_OUT) {
            ((float*)out)[lm.at(r, k)] = fmaxf(v.x * scale, clip_lo);
        } else {
            ((float2*)out)[lm.at(r, k)] = make_float2(v.x * scale, v.y * scale);
        }
    }
}

// A Fourier-space multiplier applied to every pixel of a batch of n1 x n2
// spectra (the propagator or the lens), any n1, n2.  The propagator's
// per-row factor is evaluated per element with the same float32 operations
// its column kernel tabulates.
__device__ __forceinline__ float2 prop_factor(const PropagatorMul& p, int i, int j) {
    const float fx = __fmul_rn((float)freq_int(i, p.n1), p.dfx);
    float s, c;
    fast_sincos(__fmul_rn(p.lpd, __fmul_rn(fx, fx)), &s, &c);
    return cmul(make_float2(c, -s), p.column(j));
}

template <class Mul>
__global__ __launch_bounds__(256) void k_spectrum_mul(float2* __restrict__ data, int64_t n_pix,
                                                      int n1, int n2, Mul mul) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_pix; t += stride) {
        const int j = (int)(t % n2);
        const int i = (int)((t / n2) % n1);
        if constexpr (Mul::kTable) {
            const float f2 = freq2(i, j, n1, n2, mul.dfx, mul.dfy);
            data[t] = (f2 <= mul.filter2) ? cmul(data[t], prop_factor(mul, i, j))
                                          : make_float2(0.f, 0.f);
        } else {
            data[t] = mul(data[t], i, j);
        }
    }
}

// The reference's band-limit disk index of row/column i (slice_builder_base.py
// :166-175): the entry of ifftshift(arange(-n//2, -n//2 + n)) at i, which for
// odd n is the FFT frequency minus one.
__host__ __device__ __forceinline__ int bandlimit_freq(int i, int n) {
    return (i + n / 2) % n - (n + 1) / 2;
}

// Y[s][i][j] = (sum_e X[e][s][i][j] F[e][i][j]) * D(i, j) / nPix, in element
// order, over the half spectra of a chunk of sc slices (rows ldh apart);
// written over element 0's block.  D is the reference's disk.
__global__ __launch_bounds__(256) void k_half_spectrum_sum(float2* __restrict__ X,
                                                           const float* __restrict__ sf,
                                                           int n_elems, int sc, int n1, int H,
                                                           int ldh, int n2, int bandlimit,
                                                           int band_r2, float inv_npix) {
    const int64_t total = (int64_t)sc * n1 * H;
    const int64_t plane = (int64_t)n1 * ldh;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += stride) {
        const int j = (int)(t % H);
        const int64_t si = t / H;
        const int i = (int)(si % n1);
        const int s = (int)(si / n1);
        float2 acc = make_float2(0.f, 0.f);
        for (int e = 0; e < n_elems; ++e) {
            const float2 x = X[((int64_t)e * sc + s) * plane + (int64_t)i * ldh + j];
            const float f = sf[((int64_t)e * n1 + i) * H + j];
            acc.x += x.x * f;
            acc.y += x.y * f;
        }
        float m = inv_npix;
        if (bandlimit) {
            const int ki = bandlimit_freq(i, n1), kj = bandlimit_freq(j, n2);
            if (ki * ki + kj * kj > band_r2) m = 0.f;
        }
        X[(int64_t)s * plane + (int64_t)i * ldh + j] = make_float2(acc.x * m, acc.y * m);
    }
}

}  // namespace emsim
