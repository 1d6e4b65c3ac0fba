// Wave kernels: row passes, column passes and pointwise ops for the
// multislice recurrence psi <- F^-1[ P . F[ t(V) psi ] ] / nPix
// (reference: emsim/backend/fftw_ext/src/WavePropagator.cpp:41-138).
//
// A 2-D transform is split into a row pass (contiguous axis, one N/8-thread
// group per row) and a column pass (tiles of C adjacent columns so that each
// global access is a C*8-byte contiguous segment).  The multislice step is
// fused into exactly two sweeps over the wave per slice:
//   pass A (rows):    b -> IFFT_row -> * exp(i sigma V) -> FFT_row -> a
//   pass B (columns): a -> FFT_col -> * P / nPix (masked) -> IFFT_col -> b
// so a slice costs 8+8 (pass A) + 4 (V) + 8+8 (pass B) = 36 bytes per pixel.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_lds.hpp"

namespace emsim {

// ---------------------------------------------------------------------------
// frequency helpers -- float32 arithmetic in the reference's operation order,
// never contracted into FMAs, so mask and aperture edges match exactly.
// ---------------------------------------------------------------------------

// Signed frequency of FFT bin i: the reference's ifftshift map
// is = i + (i < (n+1)/2 ? n/2 : -(n+1)/2), then is - floor(n/2)
// (WavePropagator.cpp:66-70).
__host__ __device__ __forceinline__ int freq_int(int i, int n) {
    const int is = i + (i < (n + 1) / 2 ? n / 2 : -((n + 1) / 2));
    return is - n / 2;
}

__device__ __forceinline__ float freq2(int i, int j, int n1, int n2, float dfx, float dfy) {
    const float fx = __fmul_rn((float)freq_int(i, n1), dfx);
    const float fy = __fmul_rn((float)freq_int(j, n2), dfy);
    return __fadd_rn(__fmul_rn(fx, fx), __fmul_rn(fy, fy));
}

// Fresnel propagator with the 2/3 mask and the 1/nPix of the unnormalised
// inverse (WavePropagator.cpp:62-85).  The kept set is decided exactly as
// the reference does (float32 f^2 <= filter^2); the phase factor
// exp(-i lpd f^2) is applied as exp(-i lpd fx^2) exp(-i lpd fy^2), whose
// factors each block tabulates once in LDS (prepare), so no element needs a
// sin/cos.  The split changes the float32 phase by < 1e-8 rad.
struct PropagatorMul {
    int n1, n2;
    float dfx, dfy;
    float filter2;    // (0.6667 fmax)^2
    float lpd;        // (lambda * pi) * dz
    float inv_npix;   // 1/nPix, exact for power-of-two sizes
    static constexpr bool kTable = true;

    // px[i] = exp(-i lpd fx_i^2) for all n1 rows (called by every thread of
    // the block, followed by a barrier in the caller).
    __device__ __forceinline__ void prepare(float2* px, int tid, int nthreads) const {
        for (int i = tid; i < n1; i += nthreads) {
            const float fx = __fmul_rn((float)freq_int(i, n1), dfx);
            float s, c;
            fast_sincos(__fmul_rn(lpd, __fmul_rn(fx, fx)), &s, &c);
            px[i] = make_float2(c, -s);
        }
    }
    // column factor exp(-i lpd fy_j^2) / nPix
    __device__ __forceinline__ float2 column(int j) const {
        const float fy = __fmul_rn((float)freq_int(j, n2), dfy);
        float s, c;
        fast_sincos(__fmul_rn(lpd, __fmul_rn(fy, fy)), &s, &c);
        return make_float2(c * inv_npix, -s * inv_npix);
    }
    __device__ __forceinline__ float2 apply(float2 w, int i, int j, const float2* px,
                                            float2 py) const {
        const float f2 = freq2(i, j, n1, n2, dfx, dfy);
        if (!(f2 <= filter2)) return make_float2(0.f, 0.f);
        return cmul(w, cmul(px[i], py));
    }
};

// Lens transfer function exp(-i chi) with the aperture (WavePropagator.cpp:105-135).
// chi = c1 f^4 - c2 f^2 does not factor over the axes, so each kept element
// evaluates its own sin/cos (one launch per image).
struct LensMul {
    int n1, n2;
    float dfx, dfy;
    float faper2;     // (aperture / lambda)^2
    float c1, c2;     // chi = c1 f^4 - c2 f^2
    float inv_npix;
    static constexpr bool kTable = false;
    __device__ __forceinline__ void prepare(float2*, int, int) const {}
    __device__ __forceinline__ float2 column(int) const { return make_float2(1.f, 0.f); }
    __device__ __forceinline__ float2 apply(float2 w, int i, int j, const float2*, float2) const {
        return (*this)(w, i, j);
    }
    __device__ __forceinline__ float2 operator()(float2 w, int i, int j) const {
        const float f2 = freq2(i, j, n1, n2, dfx, dfy);
        if (!(f2 <= faper2)) return make_float2(0.f, 0.f);
        const float chi = __fsub_rn(__fmul_rn(__fmul_rn(c1, f2), f2), __fmul_rn(c2, f2));
        float s, c;
        fast_sincos(chi, &s, &c);
        return make_float2((w.x * c + w.y * s) * inv_npix, (w.y * c - w.x * s) * inv_npix);
    }
};

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------
template <int N>
struct RowCfg {
    static constexpr int EPT = 8;
    static constexpr int T = N / EPT;                        // threads per row
    static constexpr int BLOCK = T >= 64 ? T : 64;           // one wave, or one row's threads
    static constexpr int G = BLOCK / T;                      // rows handled side by side
    static constexpr int WORDS = LdsLayout<N>::kWords;       // floats per row plane
};

// Column tiles: C adjacent columns per block (16 -> 128-byte contiguous
// segments per access), EPT values per thread, at most EMSIM_COL_MAX_BLOCK
// threads.
#ifndef EMSIM_COL_EPT
#define EMSIM_COL_EPT 8
#endif
#ifndef EMSIM_COL_MAX_BLOCK
#define EMSIM_COL_MAX_BLOCK 512
#endif
// minimum waves per SIMD the column kernels must leave room for (caps VGPRs)
#ifndef EMSIM_COL_MIN_WAVES
#define EMSIM_COL_MIN_WAVES 1
#endif
// SHAPE selects the tile shape (0: the built-in rule; the others are the
// alternatives the launch-shape autotuner may pick, see kColShapes):
//   1: EPT 8, half the columns per block (smaller blocks, more of them per CU)
//   2: EPT 16 (two radix-8 butterflies per thread and pass, half the threads
//      per column, the same columns per block)
template <int N, int SHAPE = 0>
struct ColCfg {
    static constexpr int EPT0 = (N / EMSIM_COL_EPT) > 256 ? N / 256
                                : (EMSIM_COL_EPT < N ? EMSIM_COL_EPT : 8);
    static constexpr int EPT = SHAPE == 2 ? 2 * EPT0 : EPT0;
    static constexpr int T = N / EPT;                        // threads per column
    static constexpr int STRIDE = LdsLayout<N>::kStride;     // odd: columns on distinct banks
    // threads per block, and 96 KiB of exchange planes (the rest of the
    // 160 KiB LDS holds the propagator's row table)
    static constexpr int CMAX_T = EMSIM_COL_MAX_BLOCK / (N / EPT0);
    static constexpr int CMAX_L = (96 * 1024) / (8 * STRIDE);
    static constexpr int CMAX = CMAX_T < CMAX_L ? CMAX_T : CMAX_L;
    static constexpr int C0 = CMAX < 1 ? 1 : (CMAX < 16 ? CMAX : 16);
    static constexpr int C = SHAPE == 1 ? (C0 > 1 ? C0 / 2 : 1) : C0;
    static constexpr int BLOCK = T * C;
};
// lengths with alternative column-tile shapes (the others use shape 0)
constexpr bool col_shapes(int n) { return n == 256 || n == 512 || n == 1024; }

// Index of this thread's transform within the block; wave-uniform when a
// transform spans whole waves, so it lives in a scalar register.
template <int T>
__device__ __forceinline__ int row_in_block() {
    const int r = threadIdx.x / T;
    if constexpr (T % 64 == 0) return __builtin_amdgcn_readfirstlane(r);
    else return r;
}

// ---------------------------------------------------------------------------
// Row worker.  A block (one wave for N <= 512) runs an Op over rpw groups of
// G rows, one after another, and issues the global loads of the next group
// before transforming the current one, so the memory latency of one row is
// hidden behind the FFTs of the previous one.  An Op provides
//   struct Regs;                                      per-thread registers
//   load(item, t, ok, Regs&)                          global -> registers
//   run(item, t, ok, Regs&, re, im, tw)               transform + store
// Every thread calls run (it contains block barriers); ok is false for
// padding items past n_items, which neither load nor store.
// ---------------------------------------------------------------------------
template <int N, class Op>
__device__ __forceinline__ void row_worker(const Op& op, int64_t n_items, int rpw,
                                           const float2* __restrict__ tw) {
    using Cfg = RowCfg<N>;
    __shared__ float lds_re[Cfg::G * Cfg::WORDS];
    __shared__ float lds_im[Cfg::G * Cfg::WORDS];
    const int gi = row_in_block<Cfg::T>();
    const int t = threadIdx.x % Cfg::T;
    float* re = lds_re + gi * Cfg::WORDS;
    float* im = lds_im + gi * Cfg::WORDS;
    const int64_t first = (int64_t)blockIdx.x * rpw * Cfg::G + gi;
    if (n_items <= 0) return;
    // Loads are never behind a branch: an item past n_items reads item
    // n_items - 1 and the last item's prefetch reads itself again; ok only
    // guards the stores.  A load behind a branch would make the compiler
    // wait for the next item's prefetch before this item's first use.
    auto clamp = [&](int64_t i) { return i < n_items ? i : n_items - 1; };
    TwCache<N, Cfg::EPT> twc;
    twc.load(tw, t);
    typename Op::Regs cur, nxt;
    op.load(clamp(first), t, true, cur);
    twc.fence();
    for (int r = 0; r < rpw; ++r) {
        const int64_t it = first + (int64_t)r * Cfg::G;
        op.load(clamp(r + 1 < rpw ? it + Cfg::G : it), t, true, nxt);
        op.run(it, t, it < n_items, cur, re, im, twc);
        cur = nxt;
    }
}

// Plain row transform of rows of length N (in place allowed).
template <int N, bool INV>
struct OpFFT {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    const float2* in;
    float2* out;
    struct Regs { float2 v[EPT]; };
    __device__ __forceinline__ void load(int64_t it, int t, bool ok, Regs& r) const {
        const float2* src = in + it * N;
#pragma unroll
        for (int m = 0; m < EPT; ++m) r.v[m] = ok ? src[t + m * T] : make_float2(0.f, 0.f);
    }
    template <class TW>
    __device__ __forceinline__ void run(int64_t it, int t, bool ok, Regs& r, float* re, float* im,
                                        const TW& tw) const {
        fft_n<N, EPT, INV>(r.v, re, im, t, tw);
        if (ok) {
            float2* dst = out + it * N;
#pragma unroll
            for (int m = 0; m < EPT; ++m) dst[t + m * T] = r.v[m];
        }
    }
};

// Multislice pass A: [IFFT_row] -> * (cos, sin)(factor V) -> FFT_row.
// FIRST: the input is the real-space wave itself (no leading IFFT).
// Rows belong to images of rows_per_img rows; the slice row of global row r
// is V + (r / rows_per_img) * v_img_stride + (r % rows_per_img) * N.
template <int N, bool FIRST>
struct OpTransmit {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    const float2* in;
    const float* V;
    float2* out;
    int rows_per_img;
    int64_t v_img_stride;
    float factor;
    struct Regs {
        float2 v[EPT];
        float p[EPT];
    };
    __device__ __forceinline__ void load(int64_t it, int t, bool ok, Regs& r) const {
        const float2* src = in + it * N;
        const int64_t img = it / rows_per_img;
        const float* vrow = V + img * v_img_stride + (it - img * rows_per_img) * N;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            r.v[m] = ok ? src[t + m * T] : make_float2(0.f, 0.f);
            r.p[m] = ok ? vrow[t + m * T] : 0.f;
        }
    }
    template <class TW>
    __device__ __forceinline__ void run(int64_t it, int t, bool ok, Regs& r, float* re, float* im,
                                        const TW& tw) const {
        if constexpr (!FIRST) fft_n<N, EPT, true>(r.v, re, im, t, tw);
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            float s, c;
            fast_sincos(__fmul_rn(r.p[m], factor), &s, &c);
            const float2 w = r.v[m];
            r.v[m] = make_float2(w.x * c - w.y * s, w.x * s + w.y * c);
        }
        fft_n<N, EPT, false>(r.v, re, im, t, tw);
        if (ok) {
            float2* dst = out + it * N;
#pragma unroll
            for (int m = 0; m < EPT; ++m) dst[t + m * T] = r.v[m];
        }
    }
};

// IFFT_row followed by |psi|^2 (the lens epilogue; the column pass already
// applied H / nPix).
template <int N>
struct OpIfftIntensity {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    const float2* in;
    float* img_out;
    struct Regs { float2 v[EPT]; };
    __device__ __forceinline__ void load(int64_t it, int t, bool ok, Regs& r) const {
        const float2* src = in + it * N;
#pragma unroll
        for (int m = 0; m < EPT; ++m) r.v[m] = ok ? src[t + m * T] : make_float2(0.f, 0.f);
    }
    template <class TW>
    __device__ __forceinline__ void run(int64_t it, int t, bool ok, Regs& r, float* re, float* im,
                                        const TW& tw) const {
        fft_n<N, EPT, true>(r.v, re, im, t, tw);
        if (ok) {
            float* dst = img_out + it * N;
#pragma unroll
            for (int m = 0; m < EPT; ++m) dst[t + m * T] = r.v[m].x * r.v[m].x + r.v[m].y * r.v[m].y;
        }
    }
};

// One named kernel per row op (distinct names in profiles).
template <int N, bool INV>
__global__ __launch_bounds__(RowCfg<N>::BLOCK) void k_row_fft(OpFFT<N, INV> op, int64_t n,
                                                               int rpw, const float2* tw) {
    row_worker<N>(op, n, rpw, tw);
}
template <int N, bool FIRST>
__global__ __launch_bounds__(RowCfg<N>::BLOCK) void k_row_transmit_fft(OpTransmit<N, FIRST> op,
                                                                        int64_t n, int rpw,
                                                                        const float2* tw) {
    row_worker<N>(op, n, rpw, tw);
}
template <int N>
__global__ __launch_bounds__(RowCfg<N>::BLOCK) void k_row_ifft_intensity(OpIfftIntensity<N> op,
                                                                          int64_t n, int rpw,
                                                                          const float2* tw) {
    row_worker<N>(op, n, rpw, tw);
}

// Rows (or row pairs) per block for the row worker: enough blocks to keep
// about 2048 waves in flight, at most 16 rows each.
inline int rows_per_block(int64_t n_groups) {
    int64_t r = n_groups / 2048;
    return (int)(r < 1 ? 1 : (r > 16 ? 16 : r));
}

// ---------------------------------------------------------------------------
// column kernel: FFT_col -> pointwise Mul(k1, j) -> IFFT_col, in place.
// n_img images of N rows x ld complex (image stride img_stride); the first
// ncols columns of each image are transformed in tiles of C columns (tiles
// numbered image-major).  Each block handles a run of tiles and loads the
// next tile while transforming the current one.
//
// Workgroups go to the 8 XCDs round-robin by linear id, so blocks L and L + 8
// share an XCD and start together.  They take the even and the odd tiles of
// the same strips of 2 C columns (128 bytes of a row when C = 8): block pair
// p = (L / 16) * 8 + L % 8 owns a run of strips and (L / 8) % 2 picks the
// parity.  Both halves of a cache line are then read while the line is in
// that XCD's L2, instead of the second half being fetched again from HBM a
// tile later.  The grid is a multiple of 16.  The strips are dealt out as
// evenly as the pair count allows (RunSched), so with a grid of two blocks
// per CU the pairs dispatched first (the larger runs) share their CUs with
// the pairs dispatched second (the smaller ones) and every CU gets about the
// same number of tiles.
// ---------------------------------------------------------------------------
constexpr unsigned kColXcds = 8;
// Work units dealt out as evenly as possible: worker u takes base + 1 units
// if u < extra, else base, starting at u base + min(u, extra).
struct RunSched {
    int base, extra;
    __device__ __forceinline__ void run(int u, int& start, int& count) const {
        count = base + (u < extra ? 1 : 0);
        start = u * base + (u < extra ? u : extra);
    }
};
__device__ __forceinline__ void col_run(unsigned L, RunSched sc, int& start, int& count,
                                        int& parity) {
    const unsigned x = L % kColXcds, k = L / kColXcds;
    sc.run((int)((k / 2) * kColXcds + x), start, count);
    parity = (int)(k % 2);
}
#ifndef EMSIM_COL_DB
#define EMSIM_COL_DB 1
#endif
constexpr bool kColDoubleBuffer = EMSIM_COL_DB != 0;

// Wave columns j in [lo, hi] that the propagator zeroes whole (none: lo >
// hi).  Column tiles skip_lo .. skip_lo + skip_n - 1 of each image hold
// only such columns; the launch runs the other live_per_img tiles.
struct DeadCols {
    int lo, hi;
    int live_per_img, skip_lo, skip_n;
    __device__ __forceinline__ bool dead(int j) const { return j >= lo && j <= hi; }
    __device__ __forceinline__ int tile(int tl) const { return tl < skip_lo ? tl : tl + skip_n; }
};

// lengths whose column transforms alternate two plane sets; their kernels
// are held to 128 VGPRs (two 8-wave blocks per CU), which the second set's
// addresses would otherwise push past
template <int N>
constexpr bool col_db() { return kColDoubleBuffer && fft_exchanges(N) % 2 == 0 && N >= 64; }

// waves per SIMD a column kernel leaves room for: 128 VGPRs with alternating
// plane sets at 8 values per thread; 16 values per thread need more
template <int N, int SHAPE>
constexpr int col_min_waves() {
    return SHAPE == 2 ? 2 : (col_db<N>() ? 4 : EMSIM_COL_MIN_WAVES);
}

template <int N, class Mul, int SHAPE = 0>
__global__ __launch_bounds__((ColCfg<N, SHAPE>::BLOCK), (col_min_waves<N, SHAPE>())) void k_col_fft_mul_ifft(
    float2* data, int ld, int ncols, DeadCols dc, int64_t img_stride, int64_t n_tiles,
    RunSched sched, Mul mul, const float2* __restrict__ tw) {
    using Cfg = ColCfg<N, SHAPE>;
    constexpr int EPT = Cfg::EPT;
    // a second plane set where the exchange count allows alternating sets
    // (one barrier per exchange instead of two)
    constexpr bool DB = col_db<N>();
    __shared__ float lds_re[Cfg::C * Cfg::STRIDE];
    __shared__ float lds_im[Cfg::C * Cfg::STRIDE];
    __shared__ float lds_re2[DB ? Cfg::C * Cfg::STRIDE : 1];
    __shared__ float lds_im2[DB ? Cfg::C * Cfg::STRIDE : 1];
    __shared__ float2 lds_px[Mul::kTable ? N : 1];
    const int c = threadIdx.x % Cfg::C;
    const int t = threadIdx.x / Cfg::C;
    float* re = lds_re + c * Cfg::STRIDE;
    float* im = lds_im + c * Cfg::STRIDE;
    float* re2 = lds_re2 + (DB ? c * Cfg::STRIDE : 0);
    float* im2 = lds_im2 + (DB ? c * Cfg::STRIDE : 0);
    if (n_tiles <= 0) return;
    mul.prepare(lds_px, threadIdx.x, Cfg::BLOCK);
    if constexpr (Mul::kTable) __syncthreads();

    int run_start, run_count, parity;
    col_run(blockIdx.x, sched, run_start, run_count, parity);
    // Tile r of this block.  Its loads are never behind a branch: a tile
    // index past n_tiles (the last blocks' tail) reads tile n_tiles - 1 and
    // a column past ncols reads column ncols - 1, and ok (tile and column
    // exist) only guards the stores.  A load behind a branch would make the
    // compiler wait for every load outstanding -- the next tile's included --
    // before the first use of this tile's values.  Tile indices are 32-bit
    // (the launch keeps n_tiles below 2^31).
    auto tile_at = [&](int r, int& col, bool& ok) __attribute__((always_inline)) -> float2* {
        const uint32_t g0 = (uint32_t)(2 * (run_start + r) + parity);
        const bool tile_ok = g0 < (uint32_t)n_tiles;
        const uint32_t g = tile_ok ? g0 : (uint32_t)n_tiles - 1u;
        const uint32_t img = g / (uint32_t)dc.live_per_img;
        col = dc.tile((int)(g - img * (uint32_t)dc.live_per_img)) * Cfg::C + c;
        ok = tile_ok && col < ncols;
        return data + (int64_t)img * img_stride + (col < ncols ? col : ncols - 1);
    };
    TwCache<N, EPT> twc;
    twc.load(tw, t);
    float2 cur[EPT], nxt[EPT];
    int col;
    bool ok;
    float2* base = tile_at(0, col, ok);
#pragma unroll
    for (int m = 0; m < EPT; ++m) cur[m] = base[(int64_t)(t + m * Cfg::T) * ld];
    twc.fence();
    for (int r = 0; r < run_count; ++r) {
        // the next tile (the last one reloads itself, so no branch)
        int col_n;
        bool ok_n;
        float2* base_n = tile_at(r + 1 < run_count ? r + 1 : r, col_n, ok_n);
#pragma unroll
        for (int m = 0; m < EPT; ++m) nxt[m] = base_n[(int64_t)(t + m * Cfg::T) * ld];
        // a dead column is zero after the propagator whatever it held
        const bool dead = dc.dead(col);
#pragma unroll
        for (int m = 0; m < EPT; ++m) cur[m] = dead ? make_float2(0.f, 0.f) : cur[m];
        const float2 py = mul.column(col);
        if constexpr (DB) fft_n_db<N, EPT, false>(cur, re, im, re2, im2, t, twc);
        else fft_n<N, EPT, false>(cur, re, im, t, twc);
#pragma unroll
        for (int m = 0; m < EPT; ++m) cur[m] = mul.apply(cur[m], t + m * Cfg::T, col, lds_px, py);
        if constexpr (DB) fft_n_db<N, EPT, true>(cur, re, im, re2, im2, t, twc);
        else fft_n<N, EPT, true>(cur, re, im, t, twc);
        if (ok) {
#pragma unroll
            for (int m = 0; m < EPT; ++m) base[(int64_t)(t + m * Cfg::T) * ld] = cur[m];
        }
#pragma unroll
        for (int m = 0; m < EPT; ++m) cur[m] = nxt[m];
        base = base_n;
        col = col_n;
        ok = ok_n;
    }
}

// ---------------------------------------------------------------------------
// pointwise kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_complex(float2* out, int64_t n, float value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = make_float2(value, 0.f);
}

__global__ void k_fill_u32(uint32_t* out, int64_t n, uint32_t value) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = value;
}

// WavePropagator.cpp:41-55: psi * (cos(factor V), sin(factor V)).
__global__ void k_transmit(const float2* in, const float* __restrict__ V, float2* out,
                           int64_t npix, int64_t n_total, int64_t v_img_stride, float factor) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_total; i += stride) {
        const int64_t img = i / npix;
        const float pix = V[img * v_img_stride + (i - img * npix)];
        float s, c;
        fast_sincos(__fmul_rn(pix, factor), &s, &c);
        const float2 w = in[i];
        out[i] = make_float2(w.x * c - w.y * s, w.x * s + w.y * c);
    }
}

__global__ void k_intensity(const float2* __restrict__ in, float* __restrict__ out, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float2 w = in[i];
        out[i] = w.x * w.x + w.y * w.y;
    }
}

}  // namespace emsim
