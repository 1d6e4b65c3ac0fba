// Potential-slice kernels: atom binning and the projected-potential build
//   V_s = C2R[ D . (1/nPix) sum_e F_e . R2C[h_{e,s}] ],  V = max(V, 1e-13)
// (reference: emsim/backend/slice_builder_base.py:166-225,
//  emsim/backend/fftw_ext/src/SliceBuilder.cpp:27-85,
//  emsim/backend/slice_builder_fftw.py:43-48).
//
// The reference does a batched R2C over all (element, slice) planes, a
// float32 sum over elements, an unnormalised C2R, and then a second
// R2C/C2R pair in numpy for the disk band-limit D.  Because D is a 0/1
// multiply on a Hermitian spectrum it commutes into the first spectrum pass,
// so here one column kernel does FFT_col -> * F_e -> sum_e -> * D / nPix ->
// IFFT_col and only one R2C and one C2R per plane remain.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"
#include "kernels_wave.hpp"

namespace emsim {

// ---------------------------------------------------------------------------
// binning
// ---------------------------------------------------------------------------
constexpr int kMaxElems = 128;

struct ElemOffsets {
    int64_t off[kMaxElems + 1];  // prefix sums of the per-element atom counts
    int n_elems;
};

// Bin of a float32 coordinate against float32 edge table E[0..nb], where
// E[b] is the float64 edge b rounded up to float32.  For float32 x,
// x >= e64 <=> x >= E[b], so this reproduces numpy.histogramdd's
// searchsorted(side='right') - 1 on the float64 edges exactly, including the
// closed last edge (only when that edge is a float32 value) and dropping NaN.
__device__ __forceinline__ int bin_of(float x, const float* __restrict__ E, int nb,
                                      int last_exact, float inv_d) {
    const float e0 = E[0];
    if (!(x >= e0)) return -1;
    const float last = E[nb];
    if (x >= last) return (x == last && last_exact) ? nb - 1 : -1;
    int b = (int)floorf(__fmul_rn(__fsub_rn(x, e0), inv_d));
    b = b < 0 ? 0 : (b > nb - 1 ? nb - 1 : b);
    while (b + 1 < nb && E[b + 1] <= x) ++b;
    while (b > 0 && E[b] > x) --b;
    return b;
}

struct BinGeom {
    const float* e0; const float* e1; const float* e2;  // edge tables (nS+1, n1+1, n2+1)
    int n0, n1, n2;
    int last0, last1, last2;
    float inv0, inv1, inv2;
    int one_slice;      // ignore column 0 (2-D histogram of columns 1, 2)
    int tile_rows;      // rows of n2 bins per LDS tile
    int tiles_per_plane;
};

// Pass 1: per atom, its tile and in-tile offset; count atoms per tile.
__global__ void k_bin_classify(const float* __restrict__ coords, int64_t n_atoms, ElemOffsets eo,
                               BinGeom g, int32_t* __restrict__ atom_tile,
                               uint32_t* __restrict__ atom_local,
                               uint32_t* __restrict__ tile_count) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_atoms; a += stride) {
        int32_t tile = -1;
        uint32_t local = 0;
        if (a < eo.off[eo.n_elems]) {
            int e = 0;
            while (e + 1 < eo.n_elems && a >= eo.off[e + 1]) ++e;
            const float z = coords[3 * a + 0], x = coords[3 * a + 1], y = coords[3 * a + 2];
            const int s = g.one_slice ? 0 : bin_of(z, g.e0, g.n0, g.last0, g.inv0);
            const int i = bin_of(x, g.e1, g.n1, g.last1, g.inv1);
            const int j = bin_of(y, g.e2, g.n2, g.last2, g.inv2);
            if (s >= 0 && i >= 0 && j >= 0) {
                const int64_t plane = (int64_t)e * g.n0 + s;
                tile = (int32_t)(plane * g.tiles_per_plane + i / g.tile_rows);
                local = (uint32_t)((i % g.tile_rows) * g.n2 + j);
                atomicAdd(&tile_count[tile], 1u);
            }
        }
        atom_tile[a] = tile;
        atom_local[a] = local;
    }
}

// Exclusive scan of n values (three-kernel scan: block sums, scan of sums,
// add).  kScanBlock threads x kScanItems values per block.
constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
    // Hillis-Steele over kScanBlock values in LDS
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const uint32_t add = tid >= off ? sh[tid - off] : 0u;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    const uint32_t incl = sh[tid];
    if (total) *total = sh[kScanBlock - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                           uint32_t* __restrict__ out, int64_t n,
                                                           uint32_t* __restrict__ block_sums) {
    __shared__ uint32_t sh[kScanBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        vals[k] = (base + k < n) ? in[base + k] : 0u;
        sum += vals[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, sh, &total);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += vals[k];
    }
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single-block exclusive scan of the block sums (n_blocks <= kScanTile * 64).
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* sums, int64_t n) {
    __shared__ uint32_t sh[kScanBlock];
    uint32_t carry = 0;
    for (int64_t start = 0; start < n; start += kScanTile) {
        const int64_t base = start + (int64_t)threadIdx.x * kScanItems;
        uint32_t vals[kScanItems];
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            vals[k] = (base + k < n) ? sums[base + k] : 0u;
            sum += vals[k];
        }
        uint32_t total;
        uint32_t run = carry + block_exclusive_scan(sum, sh, &total);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < n) sums[base + k] = run;
            run += vals[k];
        }
        carry += total;
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(uint32_t* out, int64_t n,
                                                         const uint32_t* __restrict__ block_sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    const uint32_t add = block_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) out[base + k] += add;
}

// Pass 2: scatter each binned atom's in-tile offset into its tile's bucket.
// tile_count is consumed (counted down) to hand out slots; the order inside
// a bucket does not matter because the tile histogram sums integers.
__global__ void k_bin_scatter(int64_t n_atoms, const int32_t* __restrict__ atom_tile,
                              const uint32_t* __restrict__ atom_local,
                              const uint32_t* __restrict__ tile_start, uint32_t* tile_count,
                              uint32_t* __restrict__ bucket) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_atoms; a += stride) {
        const int32_t tile = atom_tile[a];
        if (tile < 0) continue;
        const uint32_t slot = atomicSub(&tile_count[tile], 1u) - 1u;
        bucket[tile_start[tile] + slot] = atom_local[a];
    }
}

// Pass 3: one block per tile builds the tile's integer histogram with LDS
// atomics and writes it once, densely, as float32.
constexpr int kBinTileWords = 8192;  // u32 counters per LDS tile (32 KiB)
constexpr int kBinBlock = 256;

// tile_start holds n_tiles + 1 entries (the last is the number of binned atoms).
__global__ __launch_bounds__(kBinBlock) void k_bin_tiles(const uint32_t* __restrict__ bucket,
                                                         const uint32_t* __restrict__ tile_start,
                                                         BinGeom g, float* __restrict__ hist) {
    __shared__ uint32_t h[kBinTileWords];
    const int64_t tile = blockIdx.x;
    const int64_t plane = tile / g.tiles_per_plane;
    const int rb = (int)(tile - plane * g.tiles_per_plane);
    const int row0 = rb * g.tile_rows;
    const int rows = (g.n1 - row0) < g.tile_rows ? (g.n1 - row0) : g.tile_rows;
    const int words = rows * g.n2;
    for (int w = threadIdx.x; w < words; w += kBinBlock) h[w] = 0u;
    __syncthreads();
    const uint32_t beg = tile_start[tile];
    const uint32_t end = tile_start[tile + 1];
    for (uint32_t k = beg + threadIdx.x; k < end; k += kBinBlock) atomicAdd(&h[bucket[k]], 1u);
    __syncthreads();
    float* out = hist + (plane * g.n1 + row0) * (int64_t)g.n2;
    for (int w = threadIdx.x; w < words; w += kBinBlock) out[w] = (float)h[w];
}

// ---------------------------------------------------------------------------
// Water fill (slice_builder_base.py:134-147): every voxel that is empty for
// all elements receives Poisson(lam) oxygens and Poisson(2 lam) hydrogens.
// Random numbers are counter-based (SplitMix64 finaliser of seed and voxel
// index), so a result depends only on the seed, not on launch geometry.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Poisson draw by sequential inversion; exact for the small means used here.
__device__ __forceinline__ float poisson_draw(uint64_t bits, float exp_neg_lam, float lam) {
    const float u = (float)(bits >> 40) * (1.0f / 16777216.0f);
    float p = exp_neg_lam, cdf = p;
    int k = 0;
    while (u >= cdf && k < 64) {
        ++k;
        p *= lam / (float)k;
        cdf += p;
    }
    return (float)k;
}

__global__ void k_add_water(float* hist, int n_elems, int64_t vox, int idx_h, int idx_o,
                            float lam_o, float exp_o, float lam_h, float exp_h, uint64_t seed) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < vox; v += stride) {
        bool empty = true;
        for (int e = 0; e < n_elems; ++e) empty = empty && (hist[(int64_t)e * vox + v] == 0.f);
        if (!empty) continue;
        const uint64_t key = splitmix64(seed ^ splitmix64((uint64_t)v));
        const float o = poisson_draw(splitmix64(key + 1), exp_o, lam_o);
        const float h = poisson_draw(splitmix64(key + 2), exp_h, lam_h);
        hist[(int64_t)idx_o * vox + v] += o;
        hist[(int64_t)idx_h * vox + v] += h;
    }
}

// ---------------------------------------------------------------------------
// R2C of row pairs: rows a, b of length N packed as z = a + i b, one complex
// FFT, then A[k] = (Z[k] + conj Z[N-k]) / 2, B[k] = (Z[k] - conj Z[N-k]) / 2i
// for k = 0..N/2.  Plane p of the output (half rows of H = N/2+1) reads
// input plane (p / planes_per_group) * in_group_stride + p % planes_per_group
// + in_plane0 so that a chunk of slices of every element can be addressed.
// ---------------------------------------------------------------------------
template <int N>
struct OpR2CPairs {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    static constexpr int H = N / 2 + 1;
    const float* in;
    float2* out;
    int rows_per_plane, planes_per_group, ldh;
    int64_t in_group_stride, in_plane0;
    struct Regs { float2 v[EPT]; };

    __device__ __forceinline__ void locate(int64_t pair, int64_t& p, int& q) const {
        const int pairs_per_plane = rows_per_plane / 2;
        p = pair / pairs_per_plane;
        q = (int)(pair - p * pairs_per_plane);
    }
    __device__ __forceinline__ void load(int64_t pair, int t, bool ok, Regs& r) const {
        int64_t p;
        int q;
        locate(ok ? pair : 0, p, q);
        const int64_t in_plane =
            (p / planes_per_group) * in_group_stride + (p % planes_per_group) + in_plane0;
        const float* ra = in + (in_plane * rows_per_plane + 2 * q) * (int64_t)N;
        const float* rb = ra + N;
#pragma unroll
        for (int m = 0; m < EPT; ++m)
            r.v[m] = ok ? make_float2(ra[t + m * T], rb[t + m * T]) : make_float2(0.f, 0.f);
    }
    template <class TW>
    __device__ __forceinline__ void run(int64_t pair, int t, bool ok, Regs& r, float* re,
                                        float* im, const TW& tw) const {
        using L = LdsLayout<N>;
        fft_n<N, EPT, false>(r.v, re, im, t, tw);
        // full spectrum Z to LDS, then split
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int idx = L::at(t + m * T);
            re[idx] = r.v[m].x;
            im[idx] = r.v[m].y;
        }
        __syncthreads();
        if (ok) {
            int64_t p;
            int q;
            locate(pair, p, q);
            float2* oa = out + (p * rows_per_plane + 2 * q) * (int64_t)ldh;
            float2* ob = oa + ldh;
            for (int k = t; k < H; k += T) {
                const int ik = L::at(k & (N - 1));
                const int inr = L::at((N - k) & (N - 1));
                const float zr = re[ik], zi = im[ik];
                const float wr = re[inr], wi = -im[inr];   // conj Z[N-k]
                oa[k] = make_float2(0.5f * (zr + wr), 0.5f * (zi + wi));
                // (Z - conj Z[N-k]) / (2i) = -i/2 (Z - W)
                ob[k] = make_float2(0.5f * (zi - wi), -0.5f * (zr - wr));
            }
        }
        __syncthreads();   // the LDS is rewritten by the next row's transform
    }
};

// ---------------------------------------------------------------------------
// C2R of row pairs with the clip: half rows A, B (H = N/2+1) -> real rows.
// The imaginary parts at k = 0 and k = N/2 are ignored (C2R semantics), the
// rest of the spectrum is the Hermitian extension, and
// IFFT[A + iB] = a + i b.
// ---------------------------------------------------------------------------
template <int N>
struct OpC2RPairs {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    const float2* in;
    float* out;
    float clip_lo;
    int ldh;
    struct Regs { float2 v[EPT]; };
    __device__ __forceinline__ void load(int64_t pair, int t, bool ok, Regs& r) const {
        const float2* ia = in + pair * 2 * (int64_t)ldh;
        const float2* ib = ia + ldh;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int k = t + m * T;
            float2 z = make_float2(0.f, 0.f);
            if (ok) {
                const int kk = k <= N / 2 ? k : N - k;
                float2 a = ia[kk], b = ib[kk];
                if (kk == 0 || kk == N / 2) { a.y = 0.f; b.y = 0.f; }
                if (k > N / 2) { a.y = -a.y; b.y = -b.y; }
                z = make_float2(a.x - b.y, a.y + b.x);   // a + i b
            }
            r.v[m] = z;
        }
    }
    template <class TW>
    __device__ __forceinline__ void run(int64_t pair, int t, bool ok, Regs& r, float* re,
                                        float* im, const TW& tw) const {
        fft_n<N, EPT, true>(r.v, re, im, t, tw);
        if (ok) {
            float* oa = out + pair * 2 * (int64_t)N;
            float* ob = oa + N;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                oa[t + m * T] = fmaxf(r.v[m].x, clip_lo);
                ob[t + m * T] = fmaxf(r.v[m].y, clip_lo);
            }
        }
    }
};

template <int N>
__global__ __launch_bounds__(RowCfg<N>::BLOCK) void k_row_r2c_pairs(OpR2CPairs<N> op, int64_t n,
                                                                     int rpw, const float2* tw) {
    row_worker<N>(op, n, rpw, tw);
}
template <int N>
__global__ __launch_bounds__(RowCfg<N>::BLOCK) void k_row_c2r_pairs(OpC2RPairs<N> op, int64_t n,
                                                                     int rpw, const float2* tw) {
    row_worker<N>(op, n, rpw, tw);
}

// ---------------------------------------------------------------------------
// Spectrum column pass: for one chunk of n_sc slices, columns j < H of the
// half spectra X[e, sc, :, j] (e < n_elems): FFT_col, * F_e(k1, j),
// accumulate over e in element order (float32, as SliceBuilder.cpp:64-72),
// then * D(k1, j) / nPix and IFFT_col.  The result overwrites element 0's
// plane (Y[sc] = X[0, sc]); every tile reads its element-0 columns before it
// writes them, and tiles own disjoint columns.
// grid.x = column tiles, grid.y = slices of the chunk.  X rows are ldh
// complex apart (H padded to a multiple of 16, so every tile's segments are
// 64-byte aligned); the form-factor table rows are H apart.
// ---------------------------------------------------------------------------
// Columns per tile of the spectrum column kernels for length-n columns
// (SpecCfg<n>::C); the host lays out their tile-major form-factor table by it.
__host__ __device__ constexpr int spec_tile_cols(int n) {
    return (512 / (n / ((n / 8) > 256 ? n / 256 : 8))) < 1
               ? 1
               : ((512 / (n / ((n / 8) > 256 ? n / 256 : 8))) < 8
                      ? (512 / (n / ((n / 8) > 256 ? n / 256 : 8)))
                      : 8);
}

template <int N>
struct SpecCfg {
    // 8 values per thread and 8 columns per tile: 512-thread blocks, room
    // for the accumulator next to the transform without spilling
    static constexpr int EPT = (N / 8) > 256 ? N / 256 : 8;
    static constexpr int T = N / EPT;
    static constexpr int C = spec_tile_cols(N);
    static_assert(C == ((512 / T) < 1 ? 1 : ((512 / T) < 8 ? (512 / T) : 8)), "tile width");
    static constexpr int BLOCK = T * C;
    static constexpr int STRIDE = LdsLayout<N>::kStride;
};

template <int N>
__global__ __launch_bounds__(SpecCfg<N>::BLOCK) void k_col_spectrum(
    float2* X, const float* __restrict__ sf, int n_elems, int n_sc, int H, int ldh, int n2,
    int bandlimit, int band_r2, float inv_npix, const float2* __restrict__ tw) {
    using Cfg = SpecCfg<N>;
    __shared__ float lds_re[Cfg::C * Cfg::STRIDE];
    __shared__ float lds_im[Cfg::C * Cfg::STRIDE];
    const int c = threadIdx.x % Cfg::C;
    const int t = threadIdx.x / Cfg::C;
    const int col = blockIdx.x * Cfg::C + c;
    const bool valid = col < H;
    const int sc = blockIdx.y;
    float* re = lds_re + c * Cfg::STRIDE;
    float* im = lds_im + c * Cfg::STRIDE;
    const int64_t plane = (int64_t)N * ldh;
    const int64_t sf_plane = (int64_t)N * H;

    float2 acc[Cfg::EPT];
#pragma unroll
    for (int m = 0; m < Cfg::EPT; ++m) acc[m] = make_float2(0.f, 0.f);

    for (int e = 0; e < n_elems; ++e) {
        const float2* src = X + ((int64_t)e * n_sc + sc) * plane + col;
        float2 v[Cfg::EPT];
#pragma unroll
        for (int m = 0; m < Cfg::EPT; ++m)
            v[m] = valid ? src[(int64_t)(t + m * Cfg::T) * ldh] : make_float2(0.f, 0.f);
        fft_n<N, Cfg::EPT, false>(v, re, im, t, tw);
        const float* f = sf + (int64_t)e * sf_plane + col;
#pragma unroll
        for (int m = 0; m < Cfg::EPT; ++m) {
            const float fe = valid ? f[(int64_t)(t + m * Cfg::T) * H] : 0.f;
            acc[m].x += v[m].x * fe;
            acc[m].y += v[m].y * fe;
        }
    }
    // D(k1, j) / nPix: the reference's disk mask on integer frequencies
    // (slice_builder_base.py:166-175), centred as ifftshift(arange(-n//2, ...)).
    const int kj = freq_int(col, n2);
#pragma unroll
    for (int m = 0; m < Cfg::EPT; ++m) {
        const int k1 = t + m * Cfg::T;
        float s = inv_npix;
        if (bandlimit) {
            const int ki = freq_int(k1, N);
            if (ki * ki + kj * kj > band_r2) s = 0.f;
        }
        acc[m] = make_float2(acc[m].x * s, acc[m].y * s);
    }
    fft_n<N, Cfg::EPT, true>(acc, re, im, t, tw);
    if (valid) {
        float2* dst = X + (int64_t)sc * plane + col;
#pragma unroll
        for (int m = 0; m < Cfg::EPT; ++m) dst[(int64_t)(t + m * Cfg::T) * ldh] = acc[m];
    }
}

}  // namespace emsim
