// Kernels of the learned coarse-slice correction: a Fourier neural operator
// (lifting, spectral layers, projection) on channels-last activations, and
// its backward pass and optimiser.
//
// Every dense contraction is one call of k_gemm: a batched, strided GEMM
// on bf16 MFMA (v_mfma_f32_32x32x16_bf16) with float32 accumulation.
// Operands are float32 or bf16 in memory and are rounded to bf16 (nearest
// even) when staged in LDS; an optional second operand pair extends the
// sum over k (used to fuse the pointwise channel mix with the inverse
// transform of a spectral layer).  The truncated 2-D DFTs are GEMMs against
// cos/sin bases, so they run on MFMA too; only the per-mode complex channel
// mix (k_fno_mix*) and the elementwise kernels are plain VALU code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/emsim_hip.h"

namespace emsim {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

// float32 -> bf16, round to nearest even (one v_cvt_pk_bf16_f32 per pair)
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ uint32_t f32x2_to_bf16x2(float lo, float hi) {
    const f32x2_t v = {lo, hi};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2_t));
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float ld_elem(const void* p, int64_t off, int type) {
    return type == EMSIM_DT_BF16 ? bf16_to_f32(((const uint16_t*)p)[off]) : ((const float*)p)[off];
}

// erf by Abramowitz & Stegun 7.1.26 (after Hastings): branch-free, absolute
// error <= 1.5e-7, far below the bf16 rounding of the activations it feeds.
__device__ __forceinline__ float erf_as(float x) {
    const float ax = fabsf(x);
    const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.f));
    float p = fmaf(1.061405429f, t, -1.453152027f);
    p = fmaf(p, t, 1.421413741f);
    p = fmaf(p, t, -0.284496736f);
    p = fmaf(p, t, 0.254829592f);
    return copysignf(1.f - p * t * __expf(-ax * ax), x);
}

// GELU with the error function, and its derivative
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.f + erf_as(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.f + erf_as(x * 0.70710678118654752f)) +
           x * 0.39894228040143268f * __expf(-0.5f * x * x);
}

// ---------------------------------------------------------------------------
// GEMM.  Block tile 64 x 64, k staged 32 at a time; four waves, each owning a
// 32 x 32 quarter of the tile (one MFMA accumulator of 16 floats per lane).
// MFMA 32x32x16 operand layout, lane l, element j < 8:
//   A(row l % 32, k 8 (l / 32) + j),  B(k 8 (l / 32) + j, col l % 32);
// result element v < 16: D(row 8 (v / 4) + 4 (l / 32) + v % 4, col l % 32).
// ---------------------------------------------------------------------------
constexpr int kGemmM = 64, kGemmN = 64, kGemmK = 32, kGemmThreads = 256;
constexpr int kGemmLdk = kGemmK + 8;   // bf16 per LDS row: 80 bytes, 16-byte aligned rows

// Stage a 64 x 32 tile of op (rows r0.., k k0..) as bf16 rows[r][k] in LDS.
// (r, k) lives at p + r sr + k sk; outside [0, nr) x [k0, kend) reads 0.
__device__ __forceinline__ void gemm_stage(uint16_t* __restrict__ lds, const void* p, int type,
                                           int64_t sr, int64_t sk, int64_t r0, int64_t nr,
                                           int64_t k0, int64_t kend) {
    const int t = threadIdx.x;
    if (sk == 1) {
        // k contiguous: thread -> (row t / 4, 8 consecutive k)
        const int r = t >> 2, kc = (t & 3) * 8;
        const int64_t row = r0 + r;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t kk = k0 + kc + j;
            const float v = (row < nr && kk < kend) ? ld_elem(p, row * sr + kk, type) : 0.f;
            lds[r * kGemmLdk + kc + j] = f32_to_bf16(v);
        }
    } else {
        // rows contiguous (or general): thread -> (k t / 8, 8 consecutive rows)
        const int kq = t >> 3, rc = (t & 7) * 8;
        const int64_t kk = k0 + kq;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int64_t row = r0 + rc + j;
            const float v = (row < nr && kk < kend) ? ld_elem(p, row * sr + kk * sk, type) : 0.f;
            lds[(rc + j) * kGemmLdk + kq] = f32_to_bf16(v);
        }
    }
}

__device__ __forceinline__ void gemm_accumulate(f32x16_t& acc, const emsim_gemm_desc& d,
                                                const void* a, int at, int64_t a_sm, int64_t a_sk,
                                                const void* b, int bt, int64_t b_sk, int64_t b_sn,
                                                int64_t kbeg, int64_t kend, int64_t m0, int64_t n0,
                                                uint16_t* lds_a, uint16_t* lds_b) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    for (int64_t k0 = kbeg; k0 < kend; k0 += kGemmK) {
        __syncthreads();
        gemm_stage(lds_a, a, at, a_sm, a_sk, m0, d.m, k0, kend);
        // B tile stored transposed: rows = output columns, k contiguous
        gemm_stage(lds_b, b, bt, b_sn, b_sk, n0, d.n, k0, kend);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kGemmK; kk += 16) {
            const int ko = kk + 8 * (lane >> 5);
            const bf16x8_t fa = *(const bf16x8_t*)(lds_a + (wm + (lane & 31)) * kGemmLdk + ko);
            const bf16x8_t fb = *(const bf16x8_t*)(lds_b + (wn + (lane & 31)) * kGemmLdk + ko);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
    }
}

// grid: x = m tiles, y = n tiles, z = batch * splits
__global__ __launch_bounds__(kGemmThreads) void k_gemm(emsim_gemm_desc d) {
    __shared__ __align__(16) uint16_t lds_a[kGemmM * kGemmLdk];
    __shared__ __align__(16) uint16_t lds_b[kGemmN * kGemmLdk];
    const int64_t splits = d.split_k > 0 ? (d.k + d.split_k - 1) / d.split_k : 1;
    const int64_t z = blockIdx.z;
    const int64_t bi = z / splits, si = z % splits;
    const int64_t m0 = (int64_t)blockIdx.x * kGemmM, n0 = (int64_t)blockIdx.y * kGemmN;
    f32x16_t acc = {};
    const int64_t kbeg = d.split_k > 0 ? si * d.split_k : 0;
    const int64_t kend = d.split_k > 0 ? (kbeg + d.split_k < d.k ? kbeg + d.split_k : d.k) : d.k;
    const char* a = (const char*)d.a + bi * d.a_bs * (d.a_type == EMSIM_DT_BF16 ? 2 : 4);
    const char* b = (const char*)d.b + bi * d.b_bs * (d.b_type == EMSIM_DT_BF16 ? 2 : 4);
    gemm_accumulate(acc, d, a, d.a_type, d.a_sm, d.a_sk, b, d.b_type, d.b_sk, d.b_sn, kbeg, kend,
                    m0, n0, lds_a, lds_b);
    if (d.k2 > 0 && si == 0) {
        const char* a2 = (const char*)d.a2 + bi * d.a2_bs * (d.a2_type == EMSIM_DT_BF16 ? 2 : 4);
        const char* b2 = (const char*)d.b2 + bi * d.b2_bs * (d.b2_type == EMSIM_DT_BF16 ? 2 : 4);
        gemm_accumulate(acc, d, a2, d.a2_type, d.a2_sm, d.a2_sk, b2, d.b2_type, d.b2_sk, d.b2_sn,
                        0, d.k2, m0, n0, lds_a, lds_b);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t col = n0 + (w & 1) * 32 + (lane & 31);
    if (col >= d.n) return;
    const int64_t cb = d.split_k > 0 ? z * d.c_split_stride : bi * d.c_bs;
    const float bias = (d.bias && d.split_k == 0) ? d.bias[col] : 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int64_t row = m0 + (w >> 1) * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (row >= d.m) continue;
        const int64_t off = cb + row * d.c_sm + col * d.c_sn;
        float val = acc[v];
        if (d.split_k == 0) {
            if (d.beta != 0.f) val += d.beta * ld_elem(d.c, off, d.c_type);
            val += bias;
            if (d.z) {
                if (d.z_type == EMSIM_DT_BF16) ((uint16_t*)d.z)[off] = f32_to_bf16(val);
                else ((float*)d.z)[off] = val;
            }
            if (d.act == 1) val = gelu(val);
        }
        if (d.c_type == EMSIM_DT_BF16) ((uint16_t*)d.c)[off] = f32_to_bf16(val);
        else ((float*)d.c)[off] = val;
    }
}

// ---------------------------------------------------------------------------
// Fused row kernel of the 32-channel network: one block per image row
// (b, x), four waves, wave w owning the 32-pixel tiles y0 = 32 (w + 4 i).
// Channels are the MFMA rows and pixels its columns, so every operand
// fragment is 16 contiguous bytes (8 bf16 channels of one pixel, or 8
// pixels of one channel in LDS):
//   z^T = W^T v^T + Uhat^T Gy^T + b      (lifting: Lw^T X^T + b, no Uhat)
//   v'  = act(z)  -> vout (bf16), z -> zout (training), v'^T -> LDS
//   then either T^T = v'^T Fy^T, the next layer's y-transform of this row
//   (float32, (2 M2, C) per row), or the projection out = v' Qw + bq.
// Needs C = 32, k1 = 16 or 32, 2 M2 = 16 or 32, n a multiple of 32 and at
// most 512 (four 32-pixel tiles per wave).  Dynamic LDS: kFnoRowLds(n) bytes.
// ---------------------------------------------------------------------------
struct FnoRowArgs {
    const uint16_t* vin;   // (rows, n, k1) bf16
    int k1;                // input channels (C, or cin_pad for the lifting)
    const uint16_t* w16;   // (k1, 32) bf16, [k][c]
    const float* bias;     // (32)
    const uint16_t* U;     // (rows, 2 M2, 32) bf16: Uhat of every row, or null
    const uint16_t* gy;    // (n, 2 M2) bf16
    int m2x2;              // 2 M2
    int act;               // GELU after the sum
    uint16_t* vout;        // (rows, n, 32) bf16, or null
    uint16_t* zout;        // (rows, n, 32) bf16 pre-activation, or null
    const uint16_t* fy;    // (2 M2, n) bf16, or null (no y-transform of the output)
    uint16_t* T;           // (rows, 2 M2, 32) bf16
    const uint16_t* q16;   // (32, 2) bf16 projection, or null
    const float* qb;       // (2)
    float* out;            // (rows, n, 2) float32
    int n;
    // lifting straight from the simulation's outputs (mode bit kRowPack):
    // channels [Re psi / psi0, Im psi / psi0, sigma V_1 .. sigma V_ns, 0 ...]
    const float2* wave;    // (B, n, n)
    const float* slices;   // (B, ns, n, n)
    int ns;
    float inv_psi0, sigma;
    int64_t rows;          // rows of the batch (B n)
    int rpb;               // rows per block
    // mode bit kRowRelift: the lifting (16 input channels -> 32) recomputed
    // from the simulation's outputs instead of read as activations
    const uint16_t* w16_lift;   // (16, 32) bf16, [k][c]
    const float* bias_lift;     // (32)
};

constexpr int kFnoWStride = 72;    // bf16 per LDS row of W^T / Uhat^T (144 B, 16-B aligned)
__host__ __device__ constexpr int fno_vt_stride(int n) { return n + 8; }
// W^T and Uhat^T, then v'^T -- reused afterwards for the four waves'
// partial sums of T (4 x 32 x 33 float32), so it is at least that large
__host__ __device__ constexpr size_t kFnoRowLds(int n) {
    return (size_t)2 * 32 * kFnoWStride * 2 +
           ((size_t)32 * fno_vt_stride(n) * 2 > (size_t)4 * 32 * 33 * 4
                ? (size_t)32 * fno_vt_stride(n) * 2
                : (size_t)4 * 32 * 33 * 4);
}

// K1F = k1 / 16 and K2F = 2 M2 / 16 fragments per 32-pixel tile (1 or 2).
// Every global load of a wave (its tiles' input channels, Gy rows and Fy
// columns) is issued before the block's first barrier, so their latency
// overlaps the staging of W^T and Uhat^T.
// MODE bits: the sum has the Uhat term, GELU, v' written, z written, the
// next y-transform computed, the projection computed.
constexpr int kRowU = 1, kRowAct = 2, kRowV = 4, kRowZ = 8, kRowT = 16, kRowProj = 32,
              kRowPack = 64, kRowRelift = 128;
// LDS of a re-lifting row kernel beyond kFnoRowLds: the lifting's W^T
constexpr size_t kFnoReliftLds = (size_t)32 * kFnoWStride * 2;

// Lifting input channels 16 f + 8 h .. + 7 of pixel (row, y), built from the
// exit wave and potentials exactly as k_fno_pack writes them.
// Lifting input fragments of all four of a wave's tiles (K1F per tile),
// built from the exit wave and potentials exactly as k_fno_pack writes them.
// Every load is unconditional -- pixel and plane indices clamped into range,
// values selected afterwards -- so all of them are in flight at once rather
// than each waiting behind a lane-dependent or per-tile branch.
template <int K1F>
__device__ __forceinline__ void fno_packed_frags(const FnoRowArgs& a, int64_t row, int w, int l32,
                                                 int h, bf16x8_t (&vf)[4][K1F]) {
    const int n = a.n;
    const int64_t b = row / n, x = row - b * n;
    const int64_t npix = (int64_t)n * n;
    // plane 0 of this row; with no slices, any in-bounds floats (masked below)
    const float* srow = (a.ns > 0 ? a.slices + b * a.ns * npix : (const float*)a.wave) + x * n;
    const float2* wrow = a.wave + row * n;
    const int last = a.ns > 0 ? a.ns - 1 : 0;
    float raw[4][K1F][8];
    float2 wv[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int y = 32 * (w + 4 * i) + l32;
        y = y < n ? y : n - 1;
        wv[i] = wrow[y];
#pragma unroll
        for (int f = 0; f < K1F; ++f)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = 16 * f + 8 * h + e;
                const int pl = c - 2 < 0 ? 0 : (c - 2 > last ? last : c - 2);
                raw[i][f][e] = srow[pl * npix + y];
            }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool ok = 32 * (w + 4 * i) < n;
#pragma unroll
        for (int f = 0; f < K1F; ++f) {
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int c = 16 * f + 8 * h + e;
                const float val = c == 0 ? wv[i].x * a.inv_psi0
                                : c == 1 ? wv[i].y * a.inv_psi0
                                : (c < 2 + a.ns ? a.sigma * raw[i][f][e] : 0.f);
                v[e] = ok ? val : 0.f;
            }
            uint32_t q[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) q[j] = f32x2_to_bf16x2(v[2 * j], v[2 * j + 1]);
            vf[i][f] = __builtin_bit_cast(bf16x8_t, q);
        }
    }
}

// Minimum waves per SIMD the row kernel is compiled for: two (256 VGPRs,
// room for two rows' fragments without spills) for the lifting and the last
// layer; the middle layers (the Uhat term and the next y-transform), the
// 32-channel packed liftings and the training modes (z written) keep the
// compiler's choice.
__host__ __device__ constexpr int fno_rows_min_waves(int k1f, int mode) {
    return (mode & kRowRelift) ? 2
           : (((mode & kRowU) && (mode & kRowT)) || ((mode & kRowPack) && k1f > 1) ||
              (mode & kRowZ))
               ? 1
               : 2;
}

template <int K1F, int K2F, int MODE>
__global__ __launch_bounds__(256, fno_rows_min_waves(K1F, MODE)) void k_fno_rows(FnoRowArgs a) {
    constexpr bool HAS_U = MODE & kRowU, ACT = MODE & kRowAct, VOUT = MODE & kRowV,
                   ZOUT = MODE & kRowZ, TOUT = MODE & kRowT, PROJ = MODE & kRowProj,
                   RELIFT = MODE & kRowRelift;
    static_assert(!RELIFT || K1F == 2, "re-lifting feeds a 32-channel layer");
    // fragments a row loads: the layer's input, or (RELIFT) the lifting's
    constexpr int KIN = RELIFT ? 1 : K1F;
    extern __shared__ __align__(16) uint16_t fno_lds[];
    uint16_t* wT = fno_lds;                         // [c][k]
    uint16_t* uT = wT + 32 * kFnoWStride;           // [c][j]
    uint16_t* vT = uT + 32 * kFnoWStride;           // [c][y]
    constexpr int k1 = 16 * K1F, m2x2 = 16 * K2F;
    const int n = a.n, vts = fno_vt_stride(n);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int l32 = lane & 31, h = lane >> 5;
    // this block's rows; the Gy and Fy fragments and W are the same for
    // every row, so they are loaded and staged once per block
    const int64_t r_beg = (int64_t)blockIdx.x * a.rpb;
    const int64_t r_end = r_beg + a.rpb < a.rows ? r_beg + a.rpb : a.rows;

    bf16x8_t gf[4][K2F], ff[4][2];
    // Without the next y-transform (TOUT) the loads are unconditional, from
    // clamped indices and masked afterwards, so all of a wave's fragment
    // loads are in flight together instead of each tile's waiting behind its
    // branch; that measured faster for the last layer and slower for the
    // middle ones, which keep the branches.
    constexpr bool UNC = !TOUT;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int y0 = 32 * (w + 4 * i);
        const bool ok = y0 < n;
        const int y = (UNC && !ok ? 0 : y0) + l32;
#pragma unroll
        for (int f = 0; f < K2F; ++f) {
            gf[i][f] = bf16x8_t{};
            if constexpr (HAS_U && UNC) {
                const bf16x8_t t = *(const bf16x8_t*)(a.gy + (int64_t)y * m2x2 + 16 * f + 8 * h);
                gf[i][f] = ok ? t : bf16x8_t{};
            } else if constexpr (HAS_U) {
                if (ok) gf[i][f] = *(const bf16x8_t*)(a.gy + (int64_t)y * m2x2 + 16 * f + 8 * h);
            }
        }
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            ff[i][f] = bf16x8_t{};
            if (TOUT && ok && l32 < m2x2)
                ff[i][f] = *(const bf16x8_t*)(a.fy + (int64_t)l32 * n + y0 + 16 * f + 8 * h);
        }
    }

    __shared__ float lds_bias[32], lds_q[64];
    // W (k1 x 32 bf16) as four channels per thread, all loads at once
    static_assert(k1 * 8 <= 256, "one group of W per thread");
    if ((int)threadIdx.x < k1 * 8) {
        const uint2 wq = *(const uint2*)(a.w16 + 4 * threadIdx.x);
        const int k = (4 * (int)threadIdx.x) >> 5, c = (4 * (int)threadIdx.x) & 31;
        wT[c * kFnoWStride + k] = (uint16_t)wq.x;
        wT[(c + 1) * kFnoWStride + k] = (uint16_t)(wq.x >> 16);
        wT[(c + 2) * kFnoWStride + k] = (uint16_t)wq.y;
        wT[(c + 3) * kFnoWStride + k] = (uint16_t)(wq.y >> 16);
    }
    if (threadIdx.x < 32) lds_bias[threadIdx.x] = a.bias[threadIdx.x];
    if (PROJ && threadIdx.x < 64) lds_q[threadIdx.x] = bf16_to_f32(a.q16[threadIdx.x]);
    // RELIFT: the lifting's W^T (16 input channels) past the row LDS, and its bias
    __shared__ float lds_bias_lift[RELIFT ? 32 : 1];
    uint16_t* wlT = fno_lds + kFnoRowLds(n) / 2;
    if constexpr (RELIFT) {
        if ((int)threadIdx.x < 16 * 8) {
            const uint2 wq = *(const uint2*)(a.w16_lift + 4 * threadIdx.x);
            const int k = (4 * (int)threadIdx.x) >> 5, c = (4 * (int)threadIdx.x) & 31;
            wlT[c * kFnoWStride + k] = (uint16_t)wq.x;
            wlT[(c + 1) * kFnoWStride + k] = (uint16_t)(wq.x >> 16);
            wlT[(c + 2) * kFnoWStride + k] = (uint16_t)wq.y;
            wlT[(c + 3) * kFnoWStride + k] = (uint16_t)(wq.y >> 16);
        }
        if (threadIdx.x < 32) lds_bias_lift[threadIdx.x] = a.bias_lift[threadIdx.x];
    }

    // The rows are software-pipelined in pairs: the next row's input
    // fragments and Uhat row are loaded while the current row computes.
    // Every global load is unconditional (indices clamped into range), so
    // the wait before a row's first use of its fragments can leave the next
    // row's loads in flight; a load behind a branch would force a wait for
    // everything outstanding.
    constexpr int kU4 = m2x2 * 8;   // 4-channel groups of one Uhat row (2 M2 x 32 bf16)
    static_assert(kU4 <= 256, "one group of a Uhat row per thread");
    const bool u_lane = HAS_U && (int)threadIdx.x < kU4;
    const int u_grp = (int)threadIdx.x % kU4;
    // input fragments of one row; a tile beyond n reads the row's first
    // tile and is never used
    auto load_v = [&](int64_t row, bf16x8_t (&vf)[4][KIN]) __attribute__((always_inline)) {
        if constexpr (MODE & (kRowPack | kRowRelift)) {
            fno_packed_frags<KIN>(a, row, w, l32, h, vf);
        } else {
            const uint16_t* vrow = a.vin + row * n * k1;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int y0 = 32 * (w + 4 * i);
                const bool ok = y0 < n;
                const int y = (ok ? y0 : 0) + l32;
#pragma unroll
                for (int f = 0; f < K1F; ++f)
                    vf[i][f] = *(const bf16x8_t*)(vrow + (int64_t)y * k1 + 16 * f + 8 * h);
            }
        }
    };
    auto load_u = [&](int64_t row) __attribute__((always_inline)) {
        return HAS_U ? *(const uint2*)(a.U + row * m2x2 * 32 + 4 * u_grp) : make_uint2(0u, 0u);
    };
    auto body = [&](int64_t row, const bf16x8_t (&vf)[4][KIN], uint2 u_cur) __attribute__((always_inline)) {
        // the previous row is done with uT and with vT (its partial sums of T)
        __syncthreads();
        if (u_lane) {
            // channels c .. c + 3 of mode j
            const int j = (4 * (int)threadIdx.x) >> 5, c = (4 * (int)threadIdx.x) & 31;
            uT[c * kFnoWStride + j] = (uint16_t)u_cur.x;
            uT[(c + 1) * kFnoWStride + j] = (uint16_t)(u_cur.x >> 16);
            uT[(c + 2) * kFnoWStride + j] = (uint16_t)u_cur.y;
            uT[(c + 3) * kFnoWStride + j] = (uint16_t)(u_cur.y >> 16);
        }
        __syncthreads();

#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y0 = 32 * (w + 4 * i), y = y0 + l32;
            if (y0 >= n) break;
            // the layer's input fragments: as loaded, or (RELIFT) the lifting
            // recomputed exactly as the lifting kernel computes its output --
            // v = bf16(W_lift x + b_lift) -- and moved from the MFMA result
            // layout (lane: channels 8 g + 4 h + 0..3 of pixel l32) into the
            // operand layout (lane: channels 16 f + 8 h + 0..7) by swapping
            // four-channel groups with the lane 32 apart
            bf16x8_t lv[K1F];
            if constexpr (RELIFT) {
                const bf16x8_t la = *(const bf16x8_t*)(wlT + l32 * kFnoWStride + 8 * h);
                const f32x16_t zl = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la, vf[i][0], f32x16_t{}, 0, 0, 0);
                uint32_t q[4][2];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int c0 = 8 * g + 4 * h;
                    float z[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) z[j] = zl[4 * g + j] + lds_bias_lift[c0 + j];
                    q[g][0] = f32x2_to_bf16x2(z[0], z[1]);
                    q[g][1] = f32x2_to_bf16x2(z[2], z[3]);
                }
#pragma unroll
                for (int f = 0; f < K1F; ++f) {
                    const uint32_t k0 = h ? q[2 * f + 1][0] : q[2 * f][0];
                    const uint32_t k1 = h ? q[2 * f + 1][1] : q[2 * f][1];
                    const uint32_t s0 = h ? q[2 * f][0] : q[2 * f + 1][0];
                    const uint32_t s1 = h ? q[2 * f][1] : q[2 * f + 1][1];
                    const uint32_t r0 = (uint32_t)__shfl_xor((int)s0, 32);
                    const uint32_t r1 = (uint32_t)__shfl_xor((int)s1, 32);
                    uint32_t fr[4];
                    fr[0] = h ? r0 : k0;
                    fr[1] = h ? r1 : k1;
                    fr[2] = h ? k0 : r0;
                    fr[3] = h ? k1 : r1;
                    lv[f] = __builtin_bit_cast(bf16x8_t, fr);
                }
            } else {
#pragma unroll
                for (int f = 0; f < K1F; ++f) lv[f] = vf[i][f < KIN ? f : 0];
            }
            f32x16_t acc = {};
#pragma unroll
            for (int f = 0; f < K1F; ++f) {
                const bf16x8_t fa = *(const bf16x8_t*)(wT + l32 * kFnoWStride + 16 * f + 8 * h);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, lv[f], acc, 0, 0, 0);
            }
            if (HAS_U) {
#pragma unroll
                for (int f = 0; f < K2F; ++f) {
                    const bf16x8_t fa = *(const bf16x8_t*)(uT + l32 * kFnoWStride + 16 * f + 8 * h);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, gf[i][f], acc, 0, 0, 0);
                }
            }
            const int64_t pix = row * n + y;
            float p0 = 0.f, p1 = 0.f;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int c0 = 8 * g + 4 * h;
                float z[4], v[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    z[j] = acc[4 * g + j] + lds_bias[c0 + j];
                    v[j] = ACT ? gelu(z[j]) : z[j];
                }
                const uint2 vb = make_uint2(f32x2_to_bf16x2(v[0], v[1]), f32x2_to_bf16x2(v[2], v[3]));
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t word = j < 2 ? vb.x : vb.y;
                    const uint16_t bits = (uint16_t)((j & 1) ? word >> 16 : word);
                    vT[(c0 + j) * vts + y] = bits;
                    if (PROJ) {
                        const float vv = bf16_to_f32(bits);
                        p0 += vv * lds_q[2 * (c0 + j)];
                        p1 += vv * lds_q[2 * (c0 + j) + 1];
                    }
                }
                if (ZOUT)
                    *(uint2*)(a.zout + pix * 32 + c0) =
                        make_uint2(f32x2_to_bf16x2(z[0], z[1]), f32x2_to_bf16x2(z[2], z[3]));
                if (VOUT) *(uint2*)(a.vout + pix * 32 + c0) = vb;
            }
            if (PROJ) {
                p0 += __shfl_xor(p0, 32);
                p1 += __shfl_xor(p1, 32);
                if (h == 0) *(float2*)(a.out + pix * 2) = make_float2(p0 + a.qb[0], p1 + a.qb[1]);
            }
        }
        if (!TOUT) return;
        __syncthreads();
        // T^T (c, m2) = sum_y v'^T (c, y) Fy^T (y, m2), each wave over its tiles
        f32x16_t acc = {};
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int y0 = 32 * (w + 4 * i);
            if (y0 >= n) break;
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                const bf16x8_t fa = *(const bf16x8_t*)(vT + l32 * vts + y0 + 16 * f + 8 * h);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, ff[i][f], acc, 0, 0, 0);
            }
        }
        __syncthreads();   // vT is reused for the partial sums of the four waves
        float* red = (float*)vT;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int c = 8 * (v >> 2) + 4 * h + (v & 3);
            red[(w * 32 + c) * 33 + l32] = acc[v];
        }
        __syncthreads();
        // T leaves as bf16: its one consumer, the x-transform, rounds it to bf16
        uint32_t* trow = (uint32_t*)(a.T + row * m2x2 * 32);
        for (int i = threadIdx.x; i < m2x2 * 16; i += 256) {
            const int m = i >> 4, c = 2 * (i & 15);
            float s0 = 0.f, s1 = 0.f;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                s0 += red[(q * 32 + c) * 33 + m];
                s1 += red[(q * 32 + c + 1) * 33 + m];
            }
            trow[i] = f32x2_to_bf16x2(s0, s1);
        }
    };
    if (r_beg >= r_end) return;
    if constexpr (RELIFT) {
        // the raw inputs are converted as they load, so a prefetch of the
        // next row's would be waited for at once; Uhat stays one row ahead
        uint2 u_next = load_u(r_beg);
        for (int64_t row = r_beg; row < r_end; ++row) {
            bf16x8_t vf[4][KIN];
            load_v(row, vf);
            const uint2 u_cur = u_next;
            u_next = load_u(row + 1 < r_end ? row + 1 : row);
            body(row, vf, u_cur);
        }
        return;
    }
    bf16x8_t va[4][KIN], vb[4][KIN];
    load_v(r_beg, va);
    uint2 ua = load_u(r_beg), ub;
    for (int64_t row = r_beg; row < r_end; row += 2) {
        const int64_t r1 = row + 1 < r_end ? row + 1 : row;
        load_v(r1, vb);
        ub = load_u(r1);
        body(row, va, ua);
        if (row + 1 >= r_end) break;
        const int64_t r2 = row + 2 < r_end ? row + 2 : row + 1;
        load_v(r2, va);
        ua = load_u(r2);
        body(row + 1, vb, ub);
    }
}

// ---------------------------------------------------------------------------
// Spectral channel mix of one layer, one block per retained mode (k1, k2).
// PQ (batch, 2 M1, 2 M2 C): rows k1 (cos), M1 + k1 (sin) of the x-transform;
// columns (ri M2 + k2) C + c, ri = 0 / 1 the real / imaginary part of the
// y-transform.  S = sum_x exp(-i kx) T:  S_r = P_r + Q_i, S_i = P_i - Q_r.
// R (M1, M2, C, C) complex.  O = R S per mode, written to
// Ob (batch, 2 M1, M2 C): rows ri M1 + k1, columns k2 C + co.
// ---------------------------------------------------------------------------
template <int C>
__global__ __launch_bounds__(256) void k_fno_mix(const float* __restrict__ PQ,
                                                 const float2* __restrict__ R,
                                                 float* __restrict__ Ob, int batch, int M1,
                                                 int M2) {
    __shared__ float2 lds_r[C * C];
    __shared__ float2 lds_s[256 / C][C];
    const int k1 = blockIdx.x / M2, k2 = blockIdx.x % M2;
    const float2* r = R + (int64_t)blockIdx.x * C * C;
    for (int i = threadIdx.x; i < C * C; i += blockDim.x) lds_r[i] = r[i];
    const int pq_ld = 2 * M2 * C, ob_ld = M2 * C;
    const int bl = threadIdx.x / C, co = threadIdx.x % C;
    constexpr int BPB = 256 / C;   // images per pass
    for (int b0 = 0; b0 < batch; b0 += BPB) {
        __syncthreads();
        {
            const int b = b0 + bl, ci = co;
            float2 s = make_float2(0.f, 0.f);
            if (b < batch) {
                const float* pq = PQ + (int64_t)b * 2 * M1 * pq_ld;
                const float* p = pq + (int64_t)k1 * pq_ld;
                const float* q = pq + (int64_t)(M1 + k1) * pq_ld;
                const int cr = k2 * C + ci, cim = (M2 + k2) * C + ci;
                s = make_float2(p[cr] + q[cim], p[cim] - q[cr]);
            }
            lds_s[bl][ci] = s;
        }
        __syncthreads();
        const int b = b0 + bl;
        if (b < batch) {
            float orr = 0.f, oi = 0.f;
#pragma unroll 8
            for (int ci = 0; ci < C; ++ci) {
                const float2 w = lds_r[ci * C + co], s = lds_s[bl][ci];
                orr += w.x * s.x - w.y * s.y;
                oi += w.x * s.y + w.y * s.x;
            }
            float* ob = Ob + (int64_t)b * 2 * M1 * ob_ld;
            ob[(int64_t)k1 * ob_ld + k2 * C + co] = orr;
            ob[(int64_t)(M1 + k1) * ob_ld + k2 * C + co] = oi;
        }
    }
}

// ---------------------------------------------------------------------------
// Forward x-transform of a spectral layer on the fused path:
//   PQ[b] (2 M1, nc) = Fx2 (2 M1, n) T[b] (n, nc),   nc = 2 M2 C,
// with 2 M1 <= 32 MT and T in bf16.  One block per (image b, 32 V
// columns); wave w sums x in [w n/4, (w+1) n/4) into MT x V tiles of
// 32 x 32, and the four partial sums are added in LDS in two rounds.  Each
// lane loads V adjacent columns of 8 consecutive x straight from T: column
// j of the wave's V tiles is column V l32 + j of the block's slab, so the V
// values of one load are B fragments of V tiles, and a wave's load covers
// whole runs of two rows.  T is read once, with no LDS staging; Fx2
// (2 M1 n bf16) stays in cache.  Needs n % 64 == 0 and nc % (32 V) == 0.
// Same operands as k_gemm on float32 T (which it rounds to bf16); only the
// float32 summation order differs.
// ---------------------------------------------------------------------------
// V adjacent bf16 of one row, as one load
template <int V> struct Bf16V;
template <> struct Bf16V<1> { typedef uint16_t t; };
template <> struct Bf16V<2> { typedef uint32_t t; };
template <> struct Bf16V<4> { typedef uint2 t; };

// bits of element j of a Bf16V load, in the low half
template <int V>
__device__ __forceinline__ uint32_t bv_get(const typename Bf16V<V>::t& x, int j) {
    if constexpr (V == 1) return x;
    else if constexpr (V == 2) return j == 0 ? (x & 0xffffu) : (x >> 16);
    else {
        const uint32_t wd = j < 2 ? x.x : x.y;
        return (j & 1) ? (wd >> 16) : (wd & 0xffffu);
    }
}

template <int MT, int V>
__global__ __launch_bounds__(256) void k_fno_xfwd(const uint16_t* __restrict__ fx2,
                                                  const uint16_t* __restrict__ T,
                                                  float* __restrict__ PQ, int n, int m2x1,
                                                  int nc) {
    typedef typename Bf16V<V>::t fv_t;
    constexpr int TS = 32 * 33;   // one padded 32 x 32 tile in LDS
    __shared__ float red[2][MT * V * TS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l32 = lane & 31, h = lane >> 5;
    const int slabs = nc / (32 * V);
    const int64_t b = blockIdx.x / slabs;
    const int c0 = (blockIdx.x % slabs) * 32 * V;
    const int kq = n >> 2, xb = w * kq;
    bool arow[MT];
    const uint16_t* ap[MT];
#pragma unroll
    for (int i = 0; i < MT; ++i) {
        arow[i] = 32 * i + l32 < m2x1;
        ap[i] = fx2 + (int64_t)(arow[i] ? 32 * i + l32 : 0) * n + 8 * h;
    }
    const uint16_t* tp = T + b * (int64_t)n * nc + c0 + V * l32 + (int64_t)(8 * h) * nc;
    f32x16_t acc[MT][V];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < V; ++j) acc[i][j] = f32x16_t{};
    // two k-steps of 16 per iteration: every load of both is issued before
    // the first MFMA
    for (int x0 = xb; x0 < xb + kq; x0 += 32) {
        const bool two = x0 + 16 < xb + kq;   // wave-uniform
        bf16x8_t fa[2][MT];
        fv_t tv[2][8];
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int xs = s == 1 && !two ? x0 : x0 + 16 * s;
#pragma unroll
            for (int i = 0; i < MT; ++i) fa[s][i] = *(const bf16x8_t*)(ap[i] + xs);
#pragma unroll
            for (int r = 0; r < 8; ++r) tv[s][r] = *(const fv_t*)(tp + (int64_t)(xs + r) * nc);
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (s == 1 && !two) break;
#pragma unroll
            for (int j = 0; j < V; ++j) {
                uint32_t q[4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    q[r] = bv_get<V>(tv[s][2 * r], j) | (bv_get<V>(tv[s][2 * r + 1], j) << 16);
                const bf16x8_t fb = __builtin_bit_cast(bf16x8_t, q);
#pragma unroll
                for (int i = 0; i < MT; ++i) {
                    const bf16x8_t a = arow[i] ? fa[s][i] : bf16x8_t{};
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, fb, acc[i][j], 0, 0, 0);
                }
            }
        }
    }
    // partial sums: waves 2, 3 -> LDS, added by waves 0, 1; wave 1 -> LDS,
    // added by wave 0, which leaves the block's total in red[0]
    auto put = [&](float* dst) {
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    dst[(i * V + j) * TS + (8 * (v >> 2) + 4 * h + (v & 3)) * 33 + l32] = acc[i][j][v];
    };
    auto add = [&](const float* src) {
#pragma unroll
        for (int i = 0; i < MT; ++i)
#pragma unroll
            for (int j = 0; j < V; ++j)
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    acc[i][j][v] += src[(i * V + j) * TS + (8 * (v >> 2) + 4 * h + (v & 3)) * 33 + l32];
    };
    if (w >= 2) put(red[w - 2]);
    __syncthreads();
    if (w < 2) add(red[w]);
    __syncthreads();
    if (w == 1) put(red[0]);
    __syncthreads();
    if (w == 0) {
        add(red[0]);
        put(red[0]);
    }
    __syncthreads();
    float* out = PQ + b * (int64_t)m2x1 * nc + c0;
    for (int e = threadIdx.x; e < m2x1 * 32 * V; e += 256) {
        const int m = e / (32 * V), c = e % (32 * V), j = c % V, cl = c / V;
        out[(int64_t)m * nc + c] = red[0][((m >> 5) * V + j) * TS + (m & 31) * 33 + cl];
    }
}

// Inverse x-transform of a spectral layer on the fused path, written as the
// bf16 Uhat the fused row kernel reads:
//   U[b] (2 n, MC) = Gx (2 n, 2 M1) Ob[b] (2 M1, MC),   MC = M2 C.
// One block per (image b, 32 rows of U); wave w computes the 32 x 32 column
// tiles w, w + 4, ... (K = 2 M1, a multiple of 16 up to 64).  The block's
// rows are contiguous in U, so they are staged in LDS and written with
// 16-byte stores.  Rounding as k_gemm: Ob and the result to bf16, nearest
// even, with the same k order.
template <int MC>
__global__ __launch_bounds__(256) void k_fno_xinv(const uint16_t* __restrict__ gx,
                                                  const float* __restrict__ Ob,
                                                  uint16_t* __restrict__ U, int n2, int k) {
    constexpr int LD = MC + 8, TPW = MC / 128, KS = 4;
    __shared__ __align__(16) uint16_t tile[32 * LD];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l32 = lane & 31, h = lane >> 5;
    const int rtiles = n2 >> 5;
    const int64_t b = blockIdx.x / rtiles;
    const int r0 = (blockIdx.x % rtiles) * 32;
    const float* ob = Ob + b * (int64_t)k * MC + (int64_t)(8 * h) * MC;
    const int steps = k >> 4;   // 1 .. 4, wave-uniform
    bf16x8_t fa[KS];
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int sc = s < steps ? s : 0;
        fa[s] = *(const bf16x8_t*)(gx + (int64_t)(r0 + l32) * k + 16 * sc + 8 * h);
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int col = (w + 4 * t) * 32 + l32;
        float bv[KS][8];
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int sc = s < steps ? s : 0;
#pragma unroll
            for (int r = 0; r < 8; ++r) bv[s][r] = ob[(int64_t)(16 * sc + r) * MC + col];
        }
        f32x16_t acc = {};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            if (s >= steps) break;
            uint32_t q[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) q[r] = f32x2_to_bf16x2(bv[s][2 * r], bv[s][2 * r + 1]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[s], __builtin_bit_cast(bf16x8_t, q),
                                                          acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 16; ++v)
            tile[(8 * (v >> 2) + 4 * h + (v & 3)) * LD + col] = f32_to_bf16(acc[v]);
    }
    __syncthreads();
    uint16_t* ub = U + (b * n2 + r0) * (int64_t)MC;
    for (int i = threadIdx.x; i < 32 * MC / 8; i += 256) {
        const int row = i / (MC / 8), c8 = i % (MC / 8);
        *(uint4*)(ub + (int64_t)row * MC + 8 * c8) = *(const uint4*)(tile + row * LD + 8 * c8);
    }
}

// Backward of k_fno_mix: from dOb (layout of Ob) and PQ, writes dR (M1, M2,
// C, C complex, overwritten: one block per mode sums over the batch in a
// fixed order) and dPQ (layout of PQ).
template <int C>
__global__ __launch_bounds__(256) void k_fno_mix_bwd(const float* __restrict__ PQ,
                                                     const float2* __restrict__ R,
                                                     const float* __restrict__ dOb,
                                                     float2* __restrict__ dR,
                                                     float* __restrict__ dPQ, int batch, int M1,
                                                     int M2) {
    __shared__ float2 lds_r[C * C];
    __shared__ float2 lds_s[256 / C][C];
    __shared__ float2 lds_do[256 / C][C];
    const int k1 = blockIdx.x / M2, k2 = blockIdx.x % M2;
    const float2* r = R + (int64_t)blockIdx.x * C * C;
    for (int i = threadIdx.x; i < C * C; i += blockDim.x) lds_r[i] = r[i];
    const int pq_ld = 2 * M2 * C, ob_ld = M2 * C;
    constexpr int BPB = 256 / C;
    constexpr int PER = (C * C + 255) / 256;   // dR entries per thread
    float2 gr[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) gr[q] = make_float2(0.f, 0.f);
    const int bl = threadIdx.x / C, cx = threadIdx.x % C;
    for (int b0 = 0; b0 < batch; b0 += BPB) {
        __syncthreads();
        {
            const int b = b0 + bl;
            float2 s = make_float2(0.f, 0.f), g = make_float2(0.f, 0.f);
            if (b < batch) {
                const float* pq = PQ + (int64_t)b * 2 * M1 * pq_ld;
                const float* p = pq + (int64_t)k1 * pq_ld;
                const float* q = pq + (int64_t)(M1 + k1) * pq_ld;
                const int cr = k2 * C + cx, cim = (M2 + k2) * C + cx;
                s = make_float2(p[cr] + q[cim], p[cim] - q[cr]);
                const float* dob = dOb + (int64_t)b * 2 * M1 * ob_ld;
                g = make_float2(dob[(int64_t)k1 * ob_ld + k2 * C + cx],
                                dob[(int64_t)(M1 + k1) * ob_ld + k2 * C + cx]);
            }
            lds_s[bl][cx] = s;      // S(b, ci = cx)
            lds_do[bl][cx] = g;     // dO(b, co = cx)
        }
        __syncthreads();
        // dS(b, ci = cx) = sum_co conj-mix of R and dO
        const int b = b0 + bl;
        if (b < batch) {
            float sr = 0.f, si = 0.f;
#pragma unroll 8
            for (int co = 0; co < C; ++co) {
                const float2 w = lds_r[cx * C + co], g = lds_do[bl][co];
                sr += w.x * g.x + w.y * g.y;
                si += -w.y * g.x + w.x * g.y;
            }
            float* pq = dPQ + (int64_t)b * 2 * M1 * pq_ld;
            float* p = pq + (int64_t)k1 * pq_ld;
            float* q = pq + (int64_t)(M1 + k1) * pq_ld;
            const int cr = k2 * C + cx, cim = (M2 + k2) * C + cx;
            p[cr] = sr;
            q[cim] = sr;
            p[cim] = si;
            q[cr] = -si;
        }
        // dR(ci, co) += sum over this pass's images, in image order
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * 256;
            if (e < C * C) {
                const int ci = e / C, co = e % C;
                float2 acc = gr[q];
                for (int j = 0; j < BPB && b0 + j < batch; ++j) {
                    const float2 s = lds_s[j][ci], g = lds_do[j][co];
                    acc.x += s.x * g.x + s.y * g.y;
                    acc.y += -s.y * g.x + s.x * g.y;
                }
                gr[q] = acc;
            }
        }
    }
    float2* out = dR + (int64_t)blockIdx.x * C * C;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * 256;
        if (e < C * C) out[e] = gr[q];
    }
}

// dz = dv * gelu'(z) (act = 1) or dv (act = 0); z bf16, dv / dz float32
__global__ void k_fno_act_bwd(const float* __restrict__ dv, const uint16_t* __restrict__ z,
                              float* __restrict__ dz, int64_t n, int act) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dz[i] = act ? dv[i] * gelu_grad(bf16_to_f32(z[i])) : dv[i];
}

// out[j] = sum over i < rows of x[i * cols + j], two stages in a fixed
// order: k_colsum_partial writes one partial row per block, k_rowsum adds
// the partials.
__global__ __launch_bounds__(256) void k_colsum_partial(const float* __restrict__ x, int64_t rows,
                                                        int cols, float* __restrict__ part) {
    // block b sums rows [b * chunk, (b + 1) * chunk); thread (r, j) strides rows
    const int64_t chunk = (rows + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * chunk;
    const int64_t r1 = r0 + chunk < rows ? r0 + chunk : rows;
    const int lanes_per_col = 256 / cols;   // cols divides 256
    const int j = threadIdx.x % cols, rl = threadIdx.x / cols;
    float s = 0.f;
    for (int64_t r = r0 + rl; r < r1; r += lanes_per_col) s += x[r * cols + j];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if (rl == 0) {
        float t = 0.f;
        for (int q = 0; q < lanes_per_col; ++q) t += red[q * cols + j];
        part[(int64_t)blockIdx.x * cols + j] = t;
    }
}

// out[j] (+)= sum_p part[p * n + j]
__global__ void k_rowsum(const float* __restrict__ part, int np, int64_t n,
                         float* __restrict__ out, int accumulate) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        float s = 0.f;
        for (int p = 0; p < np; ++p) s += part[(int64_t)p * n + j];
        out[j] = accumulate ? out[j] + s : s;
    }
}

// Adam (bias-corrected), then the bf16 mirror of the parameters.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ p16, int64_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        p16[i] = f32_to_bf16(pi);
    }
}

__global__ void k_to_bf16(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        y[i] = f32_to_bf16(x[i]);
}

// Network input, channels-last bf16 (batch, n, n, cin_pad):
// [Re psi / psi0, Im psi / psi0, sigma V_1 .. sigma V_ns, 0 ...].
__global__ void k_fno_pack(const float2* __restrict__ wave, const float* __restrict__ slices,
                           int ns, int64_t npix, int64_t total, float inv_psi0, float sigma,
                           uint16_t* __restrict__ x, int cin_pad) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t b = i / npix, p = i - b * npix;
        const float2 w = wave[i];
        const float* v = slices + b * ns * npix + p;
        auto chan = [&](int c) -> float {
            if (c == 0) return w.x * inv_psi0;
            if (c == 1) return w.y * inv_psi0;
            return c < 2 + ns ? sigma * v[(int64_t)(c - 2) * npix] : 0.f;
        };
        // cin_pad is a multiple of 16: whole 16-byte groups of 8 channels
        uint4* xo = (uint4*)(x + i * cin_pad);
        for (int g = 0; g < cin_pad / 8; ++g) {
            uint32_t q[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                q[j] = f32_to_bf16(chan(8 * g + 2 * j)) |
                       ((uint32_t)f32_to_bf16(chan(8 * g + 2 * j + 1)) << 16);
            xo[g] = make_uint4(q[0], q[1], q[2], q[3]);
        }
    }
}

// corrected wave = coarse wave + psi0 (out_r + i out_i)
__global__ void k_fno_unpack(const float2* __restrict__ coarse, const float* __restrict__ out,
                             float psi0, float2* __restrict__ wave, int64_t total) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const float2 c = coarse[i];
        wave[i] = make_float2(c.x + psi0 * out[2 * i], c.y + psi0 * out[2 * i + 1]);
    }
}

// Per-image partial sums of |pred - target|^2, |target|^2 and |pred|^2
// (block p of image b covers pixels [p chunk, (p + 1) chunk)); part is
// (batch, nparts, 3).
__global__ __launch_bounds__(256) void k_fno_loss_partial(const float2* __restrict__ pred,
                                                          const float2* __restrict__ target,
                                                          int64_t npix, int nparts,
                                                          float* __restrict__ part) {
    const int b = blockIdx.y, pidx = blockIdx.x;
    const int64_t chunk = (npix + nparts - 1) / nparts;
    const int64_t p0 = (int64_t)pidx * chunk;
    const int64_t p1 = p0 + chunk < npix ? p0 + chunk : npix;
    float e = 0.f, t2 = 0.f, a2 = 0.f;
    for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
        const float2 a = pred[b * npix + p], t = target[b * npix + p];
        const float dx = a.x - t.x, dy = a.y - t.y;
        e += dx * dx + dy * dy;
        t2 += t.x * t.x + t.y * t.y;
        a2 += a.x * a.x + a.y * a.y;
    }
    __shared__ float re[256], rt[256], ra[256];
    re[threadIdx.x] = e;
    rt[threadIdx.x] = t2;
    ra[threadIdx.x] = a2;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            re[threadIdx.x] += re[threadIdx.x + s];
            rt[threadIdx.x] += rt[threadIdx.x + s];
            ra[threadIdx.x] += ra[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float* o = part + ((int64_t)b * nparts + pidx) * 3;
        o[0] = re[0];
        o[1] = rt[0];
        o[2] = ra[0];
    }
}

// loss (batch, 3): |pred - target|^2 / |target|^2, |target|^2, |pred|^2 per
// image; the mean of column 0 over the batch is the training loss.
__global__ void k_fno_loss_finish(const float* __restrict__ part, int nparts, int batch,
                                  float* __restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    float e = 0.f, t = 0.f, a = 0.f;
    for (int p = 0; p < nparts; ++p) {
        const float* o = part + ((int64_t)b * nparts + p) * 3;
        e += o[0];
        t += o[1];
        a += o[2];
    }
    loss[3 * b] = e / t;
    loss[3 * b + 1] = t;
    loss[3 * b + 2] = a;
}

// dout = d(mean loss) / d(network output) for pred = coarse + psi0 out:
// scale = 2 psi0 / batch.
__global__ void k_fno_loss_grad(const float2* __restrict__ pred, const float2* __restrict__ target,
                                const float* __restrict__ loss, int64_t npix, int64_t total,
                                float scale, float* __restrict__ dout) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t b = i / npix;
        const float k = scale / loss[3 * b + 1];
        const float2 a = pred[i], t = target[i];
        dout[2 * i] = k * (a.x - t.x);
        dout[2 * i + 1] = k * (a.y - t.y);
    }
}

}  // namespace emsim
