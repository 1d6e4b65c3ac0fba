// Fused batched path: atoms -> images without dense histograms or
// per-element spectra in HBM.
//
// Per batch of B molecules:
//   1. k_csr_classify / scan / k_csr_scatter: the atoms' exact bins (same
//      edge tables and rule as the histogram kernels) are sorted into rows
//      r = ((b nS + s) nE + e) n1 + x; row r's bucket holds the y bins.
// Per slice s (three launches for the whole batch):
//   2. k_spec_sparse (column tiles):  R_e(x, j) = sum_{atoms in row x} w^(j y)
//      is the row DFT of the element's histogram plane, formed directly from
//      the bucket; then FFT_col, * F_e, sum over e, * D / nPix, IFFT_col
//      -> Y (the half spectrum whose C2R is the slice's potential).
//   3. k_slice_pass_a (row pairs): C2R of two Y rows -> V rows (clip), then
//      for each of the two wave rows IFFT_row -> * exp(i sigma V) -> FFT_row.
//   4. pass B (k_col_fft_mul_ifft): FFT_col -> * P / nPix -> IFFT_col; on
//      the last slice * H P / nPix, which is the lens applied to the exit
//      wave's spectrum (F F^-1 = nPix), so the final row pass writes |psi|^2.
//
// The atom scatter fills each bucket row in whatever order its atomics
// land; k_csr_sort_rows then sorts every row's y bins, so the float32 row
// sums R_e run over the atoms in one fixed order and results are
// reproducible bit for bit (equal y bins add identical terms).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fft_lds.hpp"
#include "kernels_prep.hpp"
#include "kernels_slice.hpp"
#include "kernels_wave.hpp"

namespace emsim {

struct BatchGeom {
    BinGeom g;                 // edge tables and grid (g.n0 = nS)
    int n_elems;
    int batch;
};

// Pass 1: per atom its row and y bin; count atoms per row.  mol_off (B+1)
// and elem_off (B (nE+1)) are absolute atom offsets of molecules and of each
// molecule's element groups.  Molecules given as they come instead
// (atom_elem non-null): atom a's element index is atom_elem[a] (n_elems or
// more: not binned), elem_off is unused, and coordinates are centred on the
// fly with the molecule's bounding-box keys bbox (B, 6) -- the float32
// shift atoms.centralize applies, so the bins equal the centred host copy's.
__global__ void k_csr_classify(const float* __restrict__ coords, int64_t n_atoms,
                               const int64_t* __restrict__ mol_off,
                               const int64_t* __restrict__ elem_off, BatchGeom bg,
                               int32_t* __restrict__ atom_row, uint16_t* __restrict__ atom_y,
                               uint32_t* __restrict__ row_count,
                               const uint8_t* __restrict__ atom_elem = nullptr,
                               const uint32_t* __restrict__ bbox = nullptr) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const BinGeom& g = bg.g;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_atoms; a += stride) {
        // molecule: last b with mol_off[b] <= a
        int lo = 0, hi = bg.batch;   // mol_off[lo] <= a < mol_off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (mol_off[mid] <= a) lo = mid; else hi = mid;
        }
        const int b = lo;
        int32_t row = -1;
        uint16_t yb = 0;
        int e = -1;
        if (atom_elem) {
            const int ea = atom_elem[a];
            if (ea < bg.n_elems) e = ea;
        } else {
            const int64_t* eo = elem_off + (int64_t)b * (bg.n_elems + 1);
            if (a >= eo[0] && a < eo[bg.n_elems]) {
                e = 0;
                while (e + 1 < bg.n_elems && a >= eo[e + 1]) ++e;
            }
        }
        if (e >= 0) {
            float z = coords[3 * a + 0], x = coords[3 * a + 1], y = coords[3 * a + 2];
            if (atom_elem) {
                const uint32_t* bb = bbox + 6 * (int64_t)b;
                z = __fadd_rn(z, centring_shift(bb, 0));
                x = __fadd_rn(x, centring_shift(bb, 1));
                y = __fadd_rn(y, centring_shift(bb, 2));
            }
            // a one-slice builder bins in 2-D (x, y) and ignores z
            const int s = g.one_slice ? 0 : bin_of(z, g.e0, g.n0, g.last0, g.inv0);
            const int i = bin_of(x, g.e1, g.n1, g.last1, g.inv1);
            const int j = bin_of(y, g.e2, g.n2, g.last2, g.inv2);
            if (s >= 0 && i >= 0 && j >= 0) {
                row = (int32_t)((((int64_t)b * g.n0 + s) * bg.n_elems + e) * g.n1 + i);
                yb = (uint16_t)j;
                atomicAdd(&row_count[row], 1u);
            }
        }
        atom_row[a] = row;
        atom_y[a] = yb;
    }
}

__global__ void k_csr_scatter(int64_t n_atoms, const int32_t* __restrict__ atom_row,
                              const uint16_t* __restrict__ atom_y,
                              const uint32_t* __restrict__ row_start, uint32_t* row_count,
                              uint16_t* __restrict__ bucket) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < n_atoms; a += stride) {
        const int32_t row = atom_row[a];
        if (row < 0) continue;
        const uint32_t slot = atomicSub(&row_count[row], 1u) - 1u;
        bucket[row_start[row] + slot] = atom_y[a];
    }
}

// Sort a[0, len) ascending: insertion sort, after a gapped (Shell, Ciura's
// gaps) pre-pass for the rare long row.
__device__ __forceinline__ void sort_row(uint16_t* a, uint32_t len) {
    const uint32_t gaps[] = {701u, 301u, 132u, 57u, 23u, 10u, 4u, 1u};
    for (uint32_t gi = 0; gi < 8; ++gi) {
        const uint32_t g = gaps[gi];
        if (g >= len) continue;
        for (uint32_t i = g; i < len; ++i) {
            const uint16_t v = a[i];
            uint32_t j = i;
            while (j >= g && a[j - g] > v) {
                a[j] = a[j - g];
                j -= g;
            }
            a[j] = v;
        }
    }
}

// Sort every bucket row's y bins ascending, which fixes the order of the
// float row sums.  One block per kSortRows consecutive rows, one thread per
// row: their bucket range (contiguous) is staged in LDS when it fits, each
// thread sorts its row there, and the range is written back, so the sort's
// dependent accesses are LDS round trips; a larger range is sorted in place
// in global memory.
constexpr int kSortRows = 256;
constexpr uint32_t kSortLds = 16384;
__global__ __launch_bounds__(kSortRows) void k_csr_sort_rows(int64_t n_rows,
                                                             const uint32_t* __restrict__ row_start,
                                                             uint16_t* __restrict__ bucket) {
    __shared__ uint16_t lds[kSortLds];
    const int64_t r0 = (int64_t)blockIdx.x * kSortRows;
    const int64_t r1 = r0 + kSortRows < n_rows ? r0 + kSortRows : n_rows;
    const int64_t r = r0 + threadIdx.x;
    const uint32_t lo = row_start[r0], hi = row_start[r1];
    uint32_t beg = 0, len = 0;
    if (r < r1) {
        beg = row_start[r];
        len = row_start[r + 1] - beg;
    }
    if (hi - lo > kSortLds) {
        if (len >= 2) sort_row(bucket + beg, len);
        return;
    }
    for (uint32_t i = threadIdx.x; i < hi - lo; i += kSortRows) lds[i] = bucket[lo + i];
    __syncthreads();
    if (len >= 2) sort_row(lds + (beg - lo), len);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < hi - lo; i += kSortRows) bucket[lo + i] = lds[i];
}

// Dense count planes of one slice of every molecule, with the water fill.
// One block per (molecule b, row x): the nE rows (b, slice, e, x) of the
// atom buckets are counted into LDS (integer-valued float adds, exact and
// order-independent), then -- with water -- every voxel y empty for all
// elements gets Poisson oxygens and hydrogens drawn exactly as k_add_water
// draws them for voxel v = (slice n1 + x) n2 + y of a (n_slices, n1, n2)
// histogram with the molecule's seed; the rows are written to
// hist (n_elems, batch, n1, n2).  Dynamic LDS: n_elems * n2 floats.
__global__ __launch_bounds__(256) void k_dense_slice_rows(
    const uint32_t* __restrict__ row_start, const uint16_t* __restrict__ bucket, int slice,
    int n_slices, int n_elems, int n1, int n2, int batch, float* __restrict__ hist, int water,
    int idx_h, int idx_o, float lam_o, float exp_o, float lam_h, float exp_h,
    const uint64_t* __restrict__ seeds) {
    extern __shared__ float rows[];
    const int b = (int)(blockIdx.x / (unsigned)n1), x = (int)(blockIdx.x % (unsigned)n1);
    const int words = n_elems * n2;
    for (int i = threadIdx.x; i < words; i += blockDim.x) rows[i] = 0.f;
    __syncthreads();
    for (int e = 0; e < n_elems; ++e) {
        const int64_t r = (((int64_t)b * n_slices + slice) * n_elems + e) * n1 + x;
        const uint32_t beg = row_start[r], end = row_start[r + 1];
        for (uint32_t k = beg + threadIdx.x; k < end; k += blockDim.x)
            atomicAdd(&rows[e * n2 + bucket[k]], 1.f);
    }
    __syncthreads();
    if (water) {
        const uint64_t seed = seeds[b];
        for (int y = threadIdx.x; y < n2; y += blockDim.x) {
            bool empty = true;
            for (int e = 0; e < n_elems; ++e) empty = empty && (rows[e * n2 + y] == 0.f);
            if (!empty) continue;
            const uint64_t v = ((uint64_t)slice * (uint64_t)n1 + (uint64_t)x) * (uint64_t)n2 +
                               (uint64_t)y;
            const uint64_t key = splitmix64(seed ^ splitmix64(v));
            rows[idx_o * n2 + y] += poisson_draw(splitmix64(key + 1), exp_o, lam_o);
            rows[idx_h * n2 + y] += poisson_draw(splitmix64(key + 2), exp_h, lam_h);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < words; i += blockDim.x) {
        const int e = i / n2, y = i - e * n2;
        hist[(((int64_t)e * batch + b) * n1 + x) * n2 + y] = rows[i];
    }
}

// Spectrum-pass molecule slots of every slice: slot r of slice s is the
// molecule with the r-th most atoms in that slice (ties by index).  One
// block per slice, one thread per molecule (batch <= blockDim.x); a
// (molecule, slice)'s rows of all elements are adjacent in the row table.
__global__ void k_slice_order(const uint32_t* __restrict__ row_start, int n_slices, int n_elems,
                              int n1, int batch, int32_t* __restrict__ order) {
    __shared__ uint32_t cnt[1024];
    const int s = blockIdx.x, i = threadIdx.x;
    const int64_t rows = (int64_t)n_elems * n1;
    if (i < batch) {
        const int64_t r0 = ((int64_t)i * n_slices + s) * rows;
        cnt[i] = row_start[r0 + rows] - row_start[r0];
    }
    __syncthreads();
    if (i < batch) {
        const uint32_t c = cnt[i];
        int rank = 0;
        for (int j = 0; j < batch; ++j) {
            const uint32_t o = cnt[j];
            rank += (o > c || (o == c && j < i)) ? 1 : 0;
        }
        order[(int64_t)s * batch + rank] = i;
    }
}

// ---------------------------------------------------------------------------
// Spectrum of one slice of every molecule, from the atom buckets.
// One block per (column tile, molecule): a 1-D grid of n_tiles * batch.
// Y rows are ldh complex apart, images n1 * ldh apart.
// ---------------------------------------------------------------------------
// When the (molecule, slice)'s row table (n_elems * N + 1 entries), its atom
// y-bins and the row twiddles (n2 entries) all fit, they are staged in LDS
// once per block.  Otherwise (dense slices) the row DFTs go through the
// bucket in pieces: the rows t + m T of one m are adjacent, so their atoms
// are one contiguous bucket range, staged kSpecBucketLds entries at a time;
// the twiddles stay in LDS whenever n2 fits.  Columns of 2048 and more never
// fit the row table, so its LDS holds their longer twiddle table instead.
// Column n2/2 is folded into column 0 for N >= 2048, where the lone
// (n2/2 + 1)-th column would otherwise add a block round to single-image
// launches (1025 columns in 2-column tiles on 256 CUs).
constexpr bool spec_pack(int N) { return N >= 2048; }
constexpr int kSpecBucketLds = 8192;
// Tables are sized by N (row table for up to 8 elements, twiddles for
// n2 = N), so small transforms keep more blocks per CU; larger element counts
// or n2 > N take the piece path.
template <int N> struct SpecLds {
    static constexpr int kRows = N <= 1024 ? (8 * N + 1 < 4097 ? 8 * N + 1 : 4097) : 1;
    static constexpr int kTwq = N;
    // bucket entries staged at once; N >= 2048 runs one block per CU, which
    // leaves it most of the CU's LDS
    static constexpr int kBucket = N <= 1024 ? kSpecBucketLds : (N == 2048 ? 49152 : 24576);
};
// Element planes with at most this many atoms in the slice take the direct
// 2-D DFT instead of row DFTs + a column transform.
#ifndef EMSIM_SPEC_DIRECT
#define EMSIM_SPEC_DIRECT 4
#endif
constexpr int kSpecDirect = EMSIM_SPEC_DIRECT;
// The dispatcher hands workgroups to the 8 XCDs round-robin by linear id.
// Block L therefore runs work item spec_item(L): XCD x gets one contiguous,
// tile-major range of items, so the blocks of a column tile share that XCD's
// L2 copy of the tile's form-factor columns instead of each XCD fetching
// every tile.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned spec_item(unsigned L, unsigned total) {
    const unsigned x = L % kXcds, k = L / kXcds;
    const unsigned q = total / kXcds, rem = total % kXcds;
    return x * q + (x < rem ? x : rem) + k;
}
// minimum waves per SIMD the compiler must leave room for (4: <= 128 VGPRs,
// two 512-thread blocks per CU; larger N fit only one block in LDS anyway)
#ifndef EMSIM_SPEC_MIN_WAVES
#define EMSIM_SPEC_MIN_WAVES 4
#endif

// Copy src[0, count) to LDS: in each round a thread issues all PER of its
// loads before its stores, from indices clamped into range, so no lane's
// load waits behind a branch or behind the previous entry's store.
template <int PER, typename T>
__device__ __forceinline__ void stage_lds(T* dst, const T* __restrict__ src, uint32_t count,
                                          int tid, int nthreads) {
    for (uint32_t base = 0; base < count; base += (uint32_t)(PER * nthreads)) {
        T v[PER];
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const uint32_t i = base + (uint32_t)(tid + r * nthreads);
            v[r] = src[i < count ? i : count - 1];
        }
#pragma unroll
        for (int r = 0; r < PER; ++r) {
            const uint32_t i = base + (uint32_t)(tid + r * nthreads);
            if (i < count) dst[i] = v[r];
        }
    }
}

#ifndef EMSIM_ROW_UNROLL
#define EMSIM_ROW_UNROLL 1
#endif
// Row DFTs of this thread's EPT rows r0 + m T at column col: the twiddles
// w^(col y) of the row's atoms summed in float32, in bucket order (sorted).
// rows holds absolute bucket positions (minus base: the position of
// bkt[0]); twiddle p sits at tws[p * tw_step].
// A packed lane (column 0 carrying column n2/2) sums the real parts of
// both columns' twiddles: 1 into sr and (-1)^y into si.
// col * y < 2^24 (both below 4096), so the 24-bit multiply is exact.
template <int EPT, int T>
__device__ __forceinline__ void row_dft(float2 (&v)[EPT], bool valid, int r0, int col, int mask,
                                        const uint32_t* __restrict__ rows, uint32_t base,
                                        const uint16_t* __restrict__ bkt,
                                        const float2* __restrict__ tws, int tw_step, bool packed) {
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        float sr = 0.f, si = 0.f;
        if (valid) {
            const int r = r0 + m * T;
            const uint32_t beg = rows[r] - base, end = rows[r + 1] - base;
            if (!packed) {
                uint32_t k = beg;
#if EMSIM_ROW_UNROLL >= 2
                // two atoms per step: both twiddle loads in flight together
                for (; k + 2 <= end; k += 2) {
                    const int p0 = (int)__umul24((unsigned)col, bkt[k]) & mask;
                    const int p1 = (int)__umul24((unsigned)col, bkt[k + 1]) & mask;
                    const float2 w0 = tws[p0 * tw_step];
                    const float2 w1 = tws[p1 * tw_step];
                    sr += w0.x;
                    si += w0.y;
                    sr += w1.x;
                    si += w1.y;
                }
#endif
                for (; k < end; ++k) {
                    const float2 w = tws[((int)__umul24((unsigned)col, bkt[k]) & mask) * tw_step];
                    sr += w.x;
                    si += w.y;
                }
            } else {
                const float w0 = tws[0].x, wh = tws[((mask + 1) >> 1) * tw_step].x;
                for (uint32_t k = beg; k < end; ++k) {
                    sr += w0;
                    si += (bkt[k] & 1) ? wh : w0;
                }
            }
        }
        v[m] = make_float2(sr, si);
    }
}

// The same sums for dense slices: for each m, the bucket range of rows
// r0 - t + m T .. + T (all T threads of the column group) is staged through
// lds_bkt in pieces of kSpecBucketLds entries.  Rows come from global
// memory (rows[r] absolute), twiddles from tws[p * tw_step] (LDS when
// TWL).  Every thread of the block must call it (block barriers).
template <int EPT, int T, bool TWL>
__device__ __forceinline__ void row_dft_pieces(float2 (&v)[EPT], bool valid, int r0, int t,
                                               int col, int mask,
                                               const uint32_t* __restrict__ rows,
                                               const uint16_t* __restrict__ bucket,
                                               uint16_t* lds_bkt, const float2* tws, int tw_step,
                                               int block, bool packed) {
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int r = r0 + m * T;              // this thread's row
        const int rg = r - t;                  // first row of the group
        const uint32_t lo = __builtin_amdgcn_readfirstlane(rows[rg]);
        const uint32_t hi = __builtin_amdgcn_readfirstlane(rows[rg + T]);
        const uint32_t beg = rows[r], end = rows[r + 1];
        float sr = 0.f, si = 0.f;
        for (uint32_t p0 = lo; p0 < hi; p0 += (uint32_t)kSpecBucketLds) {
            const uint32_t p1 = hi - p0 > (uint32_t)kSpecBucketLds ? p0 + kSpecBucketLds : hi;
            __syncthreads();   // the previous piece is consumed
            for (uint32_t i = threadIdx.x; i < p1 - p0; i += block) lds_bkt[i] = bucket[p0 + i];
            __syncthreads();
            if (valid) {
                const uint32_t kb = beg > p0 ? beg : p0, ke = end < p1 ? end : p1;
                if (!packed) {
                    for (uint32_t k = kb; k < ke; ++k) {
                        const int p = (int)__umul24((unsigned)col, lds_bkt[k - p0]) & mask;
                        const float2 w = tws[p * tw_step];
                        sr += w.x;
                        si += w.y;
                    }
                } else {
                    const float w0 = tws[0].x, wh = tws[((mask + 1) >> 1) * tw_step].x;
                    for (uint32_t k = kb; k < ke; ++k) {
                        sr += w0;
                        si += (lds_bkt[k - p0] & 1) ? wh : w0;
                    }
                }
            }
        }
        v[m] = make_float2(sr, si);
    }
    __syncthreads();   // lds_bkt is free again
}

// Dense slices at N >= 2048: the element's whole bucket range
// [e_beg, e_end) is staged in pieces of CAP entries, and every thread
// walks its EPT rows' atoms in each piece (per-row accumulators), so a
// piece costs one staging and two barriers however its atoms spread over
// the rows.  Rows come from global memory (absolute positions).
template <int EPT, int T, int CAP, bool TWL>
__device__ __forceinline__ void row_dft_big(float2 (&v)[EPT], bool valid, int r0, int col,
                                            int mask, const uint32_t* __restrict__ rows,
                                            uint32_t e_beg, uint32_t e_end,
                                            const uint16_t* __restrict__ bucket,
                                            uint16_t* lds_bkt, const float2* tws, int tw_step,
                                            int block, bool packed) {
    float sr[EPT], si[EPT];
    uint32_t beg[EPT], end[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        sr[m] = 0.f;
        si[m] = 0.f;
        beg[m] = rows[r0 + m * T];
        end[m] = rows[r0 + m * T + 1];
    }
    const float w0 = tws[0].x, wh = tws[((mask + 1) >> 1) * tw_step].x;
    for (uint32_t p0 = e_beg; p0 < e_end; p0 += (uint32_t)CAP) {
        const uint32_t p1 = e_end - p0 > (uint32_t)CAP ? p0 + CAP : e_end;
        __syncthreads();   // the previous piece is consumed
        stage_lds<16>(lds_bkt, bucket + p0, p1 - p0, threadIdx.x, block);
        __syncthreads();
        if (!valid) continue;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const uint32_t kb = beg[m] > p0 ? beg[m] : p0, ke = end[m] < p1 ? end[m] : p1;
            uint32_t k = kb;
            if (!packed) {
                // four atoms per step: their loads are in flight together
                for (; k + 4 <= ke; k += 4) {
                    const uint16_t* yb = lds_bkt + (k - p0);
                    const float2 a0 = tws[((int)__umul24((unsigned)col, yb[0]) & mask) * tw_step];
                    const float2 a1 = tws[((int)__umul24((unsigned)col, yb[1]) & mask) * tw_step];
                    const float2 a2 = tws[((int)__umul24((unsigned)col, yb[2]) & mask) * tw_step];
                    const float2 a3 = tws[((int)__umul24((unsigned)col, yb[3]) & mask) * tw_step];
                    sr[m] += a0.x;
                    si[m] += a0.y;
                    sr[m] += a1.x;
                    si[m] += a1.y;
                    sr[m] += a2.x;
                    si[m] += a2.y;
                    sr[m] += a3.x;
                    si[m] += a3.y;
                }
                for (; k < ke; ++k) {
                    const float2 a =
                        tws[((int)__umul24((unsigned)col, lds_bkt[k - p0]) & mask) * tw_step];
                    sr[m] += a.x;
                    si[m] += a.y;
                }
            } else {
                for (; k < ke; ++k) {
                    sr[m] += w0;
                    si[m] += (lds_bkt[k - p0] & 1) ? wh : w0;
                }
            }
        }
    }
    __syncthreads();   // lds_bkt is free again
#pragma unroll
    for (int m = 0; m < EPT; ++m) v[m] = make_float2(sr[m], si[m]);
}

// Column transform of the spectrum kernel, twiddles from the thread's cache
// or the table.
template <int N, int EPT, bool INV, bool CACHED>
__device__ __forceinline__ void spec_fft(float2 (&v)[EPT], float* re, float* im, int t,
                                         const TwCache<N, EPT>& twc, const float2* __restrict__ tw) {
    if constexpr (CACHED) fft_n<N, EPT, INV>(v, re, im, t, twc);
    else fft_n<N, EPT, INV>(v, re, im, t, tw);
}

template <int N>
__global__ __launch_bounds__(SpecCfg<N>::BLOCK, (N >= 512 && N <= 1024) ? EMSIM_SPEC_MIN_WAVES : 1) void k_spec_sparse(
    const uint32_t* __restrict__ row_start, const uint16_t* __restrict__ bucket,
    const float* __restrict__ sf, const float* __restrict__ sft, float2* __restrict__ Y,
    int slice, int n_slices, int n_elems, int H, int ldh, int n2, int batch, int bandlimit,
    int band_r2, float inv_npix, const int* __restrict__ order, const float2* __restrict__ tw) {
    using Cfg = SpecCfg<N>;
    constexpr int EPT = Cfg::EPT;
    __shared__ float lds_re[Cfg::C * Cfg::STRIDE];
    __shared__ float lds_im[Cfg::C * Cfg::STRIDE];
    __shared__ uint32_t lds_rows[SpecLds<N>::kRows];
    __shared__ uint16_t lds_bucket[SpecLds<N>::kBucket];
    __shared__ float2 lds_twq[SpecLds<N>::kTwq];   // w^p of the row transforms (length n2)
    // (x << 16) | y of a sparse plane's atoms, in bucket order -- sorted by
    // (x, y), since rows are in x order and each row's y bins are sorted;
    // two buffers used by alternate sparse planes, so a plane's fill never
    // overwrites atoms another wave is still reading
    __shared__ uint32_t lds_atoms[2][kSpecDirect > 0 ? kSpecDirect : 1];
    // w_N^j, the x factors of the direct 2-D DFT of sparse planes (N <= 512;
    // longer columns read the table, as the LDS would then no longer hold two
    // blocks per CU, and so do their column transforms)
    constexpr bool kTwInLds = N <= 512;
    constexpr bool kTwCached = N <= 512 || N == 2048;
    __shared__ float2 lds_tw[kTwInLds ? N : 1];
    const unsigned item = spec_item(blockIdx.x, gridDim.x);
    const int tile = (int)(item / (unsigned)batch);
    // molecule slot -> molecule (null: identity).  Slots in order of the
    // slice's atom count, so each XCD's stream of blocks within a column
    // tile runs from the most expensive molecules to the cheapest.
    const int bi = (int)(item % (unsigned)batch);
    const int b = order ? __builtin_amdgcn_readfirstlane(order[(int64_t)slice * batch + bi]) : bi;
    const int c = threadIdx.x % Cfg::C;
    const int t = threadIdx.x / Cfg::C;
    const int col = tile * Cfg::C + c;
    // N >= 2048: column n2/2 rides with column 0 (spec_pack), so the tiles
    // cover columns 0 .. n2/2 - 1 only
    constexpr bool PACK = spec_pack(N);
    const int half = n2 / 2;
    const bool valid = col < (PACK ? half : H);
    const bool packed = PACK && col == 0;
    float* re = lds_re + c * Cfg::STRIDE;
    float* im = lds_im + c * Cfg::STRIDE;
    const int tw_step = kTwN / n2;
    const int64_t sf_plane = (int64_t)N * H;
    const int sf_tiles = (H + Cfg::C - 1) / Cfg::C;   // tiles of the tile-major table
    // the packed lane parks each element's transform Z_e in padding column
    // H + e of its molecule's Y rows (ldh >= H + n_elems)
    float2* zpark = Y + (int64_t)b * N * ldh + H;

    // stage this (molecule, slice)'s rows (rows of all elements are adjacent),
    // its atoms' y bins and the twiddles.  The bucket range comes straight
    // from global memory (uniform addresses, so scalar loads; readfirstlane
    // keeps the values in scalar registers and the branches on them scalar).
    // Every staging load of a thread -- row table, twiddles, the bucket's
    // first 4 x BLOCK entries -- is issued before its first LDS store, with
    // indices clamped into range rather than loads behind branches, so they
    // are in flight together and the staging costs about one round trip
    // after the bucket range; one barrier ends it.
    const int64_t rbase = (((int64_t)b * n_slices + slice) * n_elems) * N;
    const int n_rows = n_elems * N;
    const bool rows_lds = n_rows < SpecLds<N>::kRows;
    const bool tw_lds = n2 <= SpecLds<N>::kTwq;
    const uint32_t kbase = __builtin_amdgcn_readfirstlane(row_start[rbase]);
    const uint32_t kend = __builtin_amdgcn_readfirstlane(row_start[rbase + n_rows]);
    const uint32_t kcount = kend - kbase;
    // everything in LDS (the usual case), or the bucket in pieces
    const bool all_lds = rows_lds && tw_lds && kcount <= (uint32_t)SpecLds<N>::kBucket;
    constexpr int RP = (SpecLds<N>::kRows + Cfg::BLOCK - 1) / Cfg::BLOCK;
    constexpr int QP = (SpecLds<N>::kTwq + Cfg::BLOCK - 1) / Cfg::BLOCK;
    constexpr int WP = kTwInLds ? (N + Cfg::BLOCK - 1) / Cfg::BLOCK : 1;
    constexpr int BP = 4;
    uint32_t rv[RP];
    float2 qv[QP], wv[WP];
    uint16_t bv[BP];
    if (rows_lds) {
#pragma unroll
        for (int j = 0; j < RP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            rv[j] = row_start[rbase + (i <= n_rows ? i : n_rows)];
        }
    }
    if (tw_lds) {
#pragma unroll
        for (int j = 0; j < QP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            qv[j] = tw[(i < n2 ? i : n2 - 1) * tw_step];
        }
    }
    if constexpr (kTwInLds) {
#pragma unroll
        for (int j = 0; j < WP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            wv[j] = tw[(i < N ? i : N - 1) * (kTwN / N)];
        }
    }
    if (all_lds && kcount > 0) {
#pragma unroll
        for (int j = 0; j < BP; ++j) {
            const uint32_t i = threadIdx.x + (uint32_t)(j * Cfg::BLOCK);
            bv[j] = bucket[kbase + (i < kcount ? i : kcount - 1)];
        }
    }
    // the column transforms' twiddles in registers (28 VGPRs at N = 512, 42
    // at 2048, where one block per CU leaves room; the LDS table's reads
    // shared banks, the global table's are L2 round trips); N = 1024 and
    // 4096 read the table, as their register budgets have no room
    TwCache<N, EPT> twc;
    if constexpr (kTwCached) twc.load(tw, t);
    if (rows_lds) {
#pragma unroll
        for (int j = 0; j < RP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            if (i <= n_rows) lds_rows[i] = rv[j];
        }
    }
    if (tw_lds) {
#pragma unroll
        for (int j = 0; j < QP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            if (i < n2) lds_twq[i] = qv[j];
        }
    }
    if constexpr (kTwInLds) {
#pragma unroll
        for (int j = 0; j < WP; ++j) {
            const int i = (int)threadIdx.x + j * Cfg::BLOCK;
            if (i < N) lds_tw[i] = wv[j];
        }
    }
    if (all_lds && kcount > 0) {
#pragma unroll
        for (int j = 0; j < BP; ++j) {
            const uint32_t i = threadIdx.x + (uint32_t)(j * Cfg::BLOCK);
            if (i < kcount) lds_bucket[i] = bv[j];
        }
        // the rest of a long bucket (more than BP x BLOCK atoms)
        const uint32_t done = (uint32_t)(BP * Cfg::BLOCK);
        if (kcount > done)
            stage_lds<BP>(lds_bucket + done, bucket + kbase + done, kcount - done, threadIdx.x,
                          Cfg::BLOCK);
    }
    __syncthreads();

    float2 acc[EPT];
#pragma unroll
    for (int m = 0; m < EPT; ++m) acc[m] = make_float2(0.f, 0.f);

    bool any = false;   // block-uniform: some element has atoms in this slice
    unsigned n_direct = 0;   // sparse planes so far (block-uniform)
    for (int e = 0; e < n_elems; ++e) {
        uint32_t e_beg, e_end;
        if (rows_lds) {
            e_beg = __builtin_amdgcn_readfirstlane(lds_rows[e * N]);
            e_end = __builtin_amdgcn_readfirstlane(lds_rows[(e + 1) * N]);
        } else {
            e_beg = __builtin_amdgcn_readfirstlane(row_start[rbase + e * N]);
            e_end = __builtin_amdgcn_readfirstlane(row_start[rbase + (e + 1) * N]);
        }
        const uint32_t n_e = e_end - e_beg;
        // an element without atoms in this slice adds exactly zero
        if (n_e == 0) {
            if (packed) {
#pragma unroll
                for (int m = 0; m < EPT; ++m)
                    zpark[(int64_t)(t + m * Cfg::T) * ldh + e] = make_float2(0.f, 0.f);
            }
            continue;
        }
        any = true;
        // form factors first: their loads stay in flight behind the transform.
        // The tile-major table holds this tile's C columns side by side in
        // every row (zeros beyond H), so a wave's load is one contiguous run.
        const float* f = sft + (((int64_t)e * sf_tiles + tile) * N + t) * Cfg::C + c;
        float fe[EPT];
#pragma unroll
        for (int m = 0; m < EPT; ++m) fe[m] = f[(int64_t)m * Cfg::T * Cfg::C];
        float2 v[EPT];
        if (all_lds && n_e <= (uint32_t)kSpecDirect) {
            // few atoms: their 2-D DFT sum_a w_N^(k x_a) w_n2^(col y_a) directly,
            // cheaper than a column transform (the y factor is the twiddle the
            // row DFT would use, the x factor the transform's table entry)
            uint32_t* atoms = lds_atoms[n_direct & 1];
            ++n_direct;
            if (c == 0) {
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int x = t + m * Cfg::T;
                    const uint32_t beg = lds_rows[e * N + x], end = lds_rows[e * N + x + 1];
                    for (uint32_t k = beg; k < end; ++k)
                        atoms[k - e_beg] = ((uint32_t)x << 16) | lds_bucket[k - kbase];
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < EPT; ++m) v[m] = make_float2(0.f, 0.f);
            // in (x, y) order: a fixed order for the float sum
            for (uint32_t a = 0; a < n_e; ++a) {
                const uint32_t xy = atoms[a];
                const int x = (int)(xy >> 16), y = (int)(xy & 0xffffu);
                float2 wy = lds_twq[(col * y) & (n2 - 1)];
                if (packed)   // column 0 (1) + i column n2/2 ((-1)^y), both real
                    wy = make_float2(lds_twq[0].x, (y & 1) ? lds_twq[half].x : lds_twq[0].x);
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int p = ((t + m * Cfg::T) * x) & (N - 1);
                    float2 wx;
                    if constexpr (kTwInLds) wx = lds_tw[p];
                    else wx = tw[p * (kTwN / N)];
                    v[m].x += wx.x * wy.x - wx.y * wy.y;
                    v[m].y += wx.x * wy.y + wx.y * wy.x;
                }
            }
        } else {
            // block-uniform choice, so each path's loads have one address space
            if (all_lds)
                row_dft<EPT, Cfg::T>(v, valid, e * N + t, col, n2 - 1, lds_rows, kbase, lds_bucket,
                                     lds_twq, 1, packed);
            else if constexpr (N >= 2048) {
                if (tw_lds)
                    row_dft_big<EPT, Cfg::T, SpecLds<N>::kBucket, true>(
                        v, valid, e * N + t, col, n2 - 1, row_start + rbase, e_beg, e_end, bucket,
                        lds_bucket, lds_twq, 1, Cfg::BLOCK, packed);
                else
                    row_dft_big<EPT, Cfg::T, SpecLds<N>::kBucket, false>(
                        v, valid, e * N + t, col, n2 - 1, row_start + rbase, e_beg, e_end, bucket,
                        lds_bucket, tw, tw_step, Cfg::BLOCK, packed);
            } else if (tw_lds)
                row_dft_pieces<EPT, Cfg::T, true>(v, valid, e * N + t, t, col, n2 - 1,
                                                  row_start + rbase, bucket, lds_bucket, lds_twq,
                                                  1, Cfg::BLOCK, packed);
            else
                row_dft_pieces<EPT, Cfg::T, false>(v, valid, e * N + t, t, col, n2 - 1,
                                                   row_start + rbase, bucket, lds_bucket, tw,
                                                   tw_step, Cfg::BLOCK, packed);
            spec_fft<N, EPT, false, kTwCached>(v, re, im, t, twc, tw);
        }
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            acc[m].x += v[m].x * fe[m];
            acc[m].y += v[m].y * fe[m];
        }
        if (packed) {
#pragma unroll
            for (int m = 0; m < EPT; ++m) zpark[(int64_t)(t + m * Cfg::T) * ldh + e] = v[m];
        }
    }
    if (!any) {
        // empty slice of this molecule: the half spectrum is exactly zero
        if (valid) {
            float2* dst = Y + (int64_t)b * N * ldh + col;
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                dst[(int64_t)(t + m * Cfg::T) * ldh] = make_float2(0.f, 0.f);
                if (packed) dst[(int64_t)(t + m * Cfg::T) * ldh + half] = make_float2(0.f, 0.f);
            }
        }
        return;
    }
    const int kj = freq_int(col, n2);
#pragma unroll
    for (int m = 0; m < EPT; ++m) {
        const int k1 = t + m * Cfg::T;
        float s = inv_npix;
        if (bandlimit) {
            const int ki = freq_int(k1, N);
            if (ki * ki + kj * kj > band_r2) s = 0.f;
        }
        acc[m] = make_float2(acc[m].x * s, acc[m].y * s);
    }
    if (PACK && tile == 0) {
        // Column n2/2 from column 0's packed transforms Z_e = F0_e + i FN_e
        // (both columns' row DFTs are real, the form factors and masks even
        // in kx): P = sum_e sf(k, 0) Z_e (acc, already scaled) and
        // Q = sum_e sf(k, n2/2) Z_e give A0 = (P(k) + conj P(-k)) / 2 and
        // AN = (Q(k) - conj Q(-k)) / 2i; one inverse transform of A0 + i AN
        // returns both real columns.  Block-uniform branch: every thread
        // reaches the barriers.
        using L = LdsLayout<N>;
        float* re1 = lds_re + Cfg::STRIDE;   // column 1's planes, free here
        float* im1 = lds_im + Cfg::STRIDE;
        if (packed) {
            float2 qa[EPT];
#pragma unroll
            for (int m = 0; m < EPT; ++m) qa[m] = make_float2(0.f, 0.f);
            for (int e = 0; e < n_elems; ++e) {   // element-outer: EPT loads in flight
                const float* fh = sf + (int64_t)e * sf_plane + half;
#pragma unroll
                for (int m = 0; m < EPT; ++m) {
                    const int64_t k1 = t + m * Cfg::T;
                    const float2 ze = zpark[k1 * ldh + e];
                    const float fq = fh[k1 * H];
                    qa[m].x += ze.x * fq;
                    qa[m].y += ze.y * fq;
                }
            }
            const int kh = freq_int(half, n2);
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int k1 = t + m * Cfg::T;
                float s = inv_npix;
                if (bandlimit) {
                    const int ki = freq_int(k1, N);
                    if (ki * ki + kh * kh > band_r2) s = 0.f;
                }
                re[L::at(k1)] = acc[m].x;
                im[L::at(k1)] = acc[m].y;
                re1[L::at(k1)] = qa[m].x * s;
                im1[L::at(k1)] = qa[m].y * s;
            }
        }
        __syncthreads();
        if (packed) {
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int k1 = t + m * Cfg::T;
                const int nk = (N - k1) & (N - 1);
                const float2 p = acc[m];
                const float2 q = make_float2(re1[L::at(k1)], im1[L::at(k1)]);
                const float2 pm = make_float2(re[L::at(nk)], im[L::at(nk)]);
                const float2 qm = make_float2(re1[L::at(nk)], im1[L::at(nk)]);
                const float2 a0 = make_float2(0.5f * (p.x + pm.x), 0.5f * (p.y - pm.y));
                const float2 an = make_float2(0.5f * (q.y + qm.y), -0.5f * (q.x - qm.x));
                acc[m] = make_float2(a0.x - an.y, a0.y + an.x);
            }
        }
        __syncthreads();
    }
    spec_fft<N, EPT, true, kTwCached>(acc, re, im, t, twc, tw);
    if (valid) {
        float2* dst = Y + (int64_t)b * N * ldh + col;
        if (!packed) {
#pragma unroll
            for (int m = 0; m < EPT; ++m) dst[(int64_t)(t + m * Cfg::T) * ldh] = acc[m];
        } else {
            // y0 + i yN: the imaginary parts of columns 0 and n2/2 are zero
#pragma unroll
            for (int m = 0; m < EPT; ++m) {
                const int64_t o = (int64_t)(t + m * Cfg::T) * ldh;
                dst[o] = make_float2(acc[m].x, 0.f);
                dst[o + half] = make_float2(acc[m].y, 0.f);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Pass A of one slice with the potential's C2R fused in: item = row pair
// (image b, rows 2q, 2q+1).  FIRST: the incoming wave is the constant psi0.
// ---------------------------------------------------------------------------
template <int N, bool FIRST>
struct OpSlicePassA {
    static constexpr int T = RowCfg<N>::T, EPT = RowCfg<N>::EPT;
    const float2* Y;       // (B, n1, ldh) half spectra of this slice
    float2* wave;          // (B, n1, N), read (unless FIRST) and written in place
    int ldh;
    float psi0;
    float factor;          // lambda * gamma
    float* vout;           // nullable: the slice's potential rows, (B, nS, n1, N)
    int64_t v_img_stride;  // nS n1 N
    int64_t v_slice_off;   // slice n1 N
    int rows_per_img;      // n1
    // wave columns [dead_lo, dead_hi] are zero after every pass B (the
    // propagator's mask); the buffer holds zeros there from the start of
    // the run, and this pass never stores into them, so they are loaded as
    // they are
    int dead_lo, dead_hi;
    __device__ __forceinline__ bool dead(int k) const { return k >= dead_lo && k <= dead_hi; }
    // raw half-spectrum values of the pair's two rows at this thread's k
    __device__ __forceinline__ void load_y(int64_t pair, int t, bool ok, float2 (&ya)[EPT],
                                           float2 (&yb)[EPT]) const {
        const float2* ia = Y + pair * 2 * (int64_t)ldh;
        const float2* ib = ia + ldh;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int k = t + m * T;
            const int kk = k <= N / 2 ? k : N - k;
            ya[m] = ok ? ia[kk] : make_float2(0.f, 0.f);
            yb[m] = ok ? ib[kk] : make_float2(0.f, 0.f);
        }
    }
    __device__ __forceinline__ void load_w(int64_t pair, int row, int t, bool ok,
                                           float2 (&w)[EPT]) const {
        const float2* src = wave + (pair * 2 + row) * (int64_t)N;
#pragma unroll
        for (int m = 0; m < EPT; ++m) w[m] = ok ? src[t + m * T] : make_float2(0.f, 0.f);
    }
    // Hermitian extension of both rows, packed as a + i b for one complex
    // inverse FFT whose real / imaginary parts are the two rows' C2R.
    __device__ __forceinline__ void pack(const float2 (&ya)[EPT], const float2 (&yb)[EPT], int t,
                                         float2 (&z)[EPT]) const {
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            const int k = t + m * T;
            float2 a = ya[m], b = yb[m];
            if (k == 0 || k == N / 2) { a.y = 0.f; b.y = 0.f; }
            if (k > N / 2) { a.y = -a.y; b.y = -b.y; }
            z[m] = make_float2(a.x - b.y, a.y + b.x);
        }
    }
    template <class TW>
    __device__ __forceinline__ void transmit_row(float2 (&w)[EPT], const float (&V)[EPT], int t,
                                                 float* re, float* im, const TW& tw) const {
        if constexpr (FIRST) {
#pragma unroll
            for (int m = 0; m < EPT; ++m) w[m] = make_float2(psi0, 0.f);
        } else {
            fft_n<N, EPT, true>(w, re, im, t, tw);
        }
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            float s, c;
            fast_sincos(__fmul_rn(V[m], factor), &s, &c);
            const float2 x = w[m];
            w[m] = make_float2(x.x * c - x.y * s, x.x * s + x.y * c);
        }
        fft_n<N, EPT, false>(w, re, im, t, tw);
    }
    __device__ __forceinline__ void store_v(int64_t pair, int t, bool ok, const float (&va)[EPT],
                                            const float (&vb)[EPT]) const {
        if (!ok || !vout) return;
        const int64_t r = pair * 2, img = r / rows_per_img;
        float* dst = vout + img * v_img_stride + v_slice_off + (r - img * rows_per_img) * N;
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            dst[t + m * T] = va[m];
            dst[N + t + m * T] = vb[m];
        }
    }
    // bit m: this thread's element m is a live column (the same for every row)
    __device__ __forceinline__ uint32_t live_mask(int t) const {
        uint32_t bits = 0;
#pragma unroll
        for (int m = 0; m < EPT; ++m) bits |= dead(t + m * T) ? 0u : (1u << m);
        return bits;
    }
    __device__ __forceinline__ void store_w(int64_t pair, int row, int t, bool ok, uint32_t live,
                                            const float2 (&w)[EPT]) const {
        if (!ok) return;
        float2* dst = wave + (pair * 2 + row) * (int64_t)N;
#pragma unroll
        for (int m = 0; m < EPT; ++m)
            if (live & (1u << m)) dst[t + m * T] = w[m];
    }
};

// minimum waves per SIMD pass A must leave room for (caps VGPRs)
#ifndef EMSIM_PASSA_MIN_WAVES
#define EMSIM_PASSA_MIN_WAVES 1
#endif
// A block (one wave for N <= 512) handles a run of groups of G row pairs
// (RunSched: runs of base or base + 1 groups).  The
// loads are staged so that each is in flight behind transforms: row 1's wave
// during the C2R and row 0's transforms, and the next pair's half spectra
// and row 0 during row 1's transforms.
template <int N, bool FIRST>
__global__ __launch_bounds__(RowCfg<N>::BLOCK, EMSIM_PASSA_MIN_WAVES) void k_slice_pass_a(
    OpSlicePassA<N, FIRST> op, int64_t n_pairs, RunSched sched, const float2* tw) {
    using Cfg = RowCfg<N>;
    constexpr int EPT = Cfg::EPT;
    __shared__ float lds_re[Cfg::G * Cfg::WORDS];
    __shared__ float lds_im[Cfg::G * Cfg::WORDS];
    const int gi = row_in_block<Cfg::T>();
    const int t = threadIdx.x % Cfg::T;
    float* re = lds_re + gi * Cfg::WORDS;
    float* im = lds_im + gi * Cfg::WORDS;
    // this block's run of row-pair groups (G pairs side by side)
    int run_start, run_count;
    sched.run((int)blockIdx.x, run_start, run_count);
    const int64_t first = (int64_t)run_start * Cfg::G + gi;
    if (n_pairs <= 0) return;
    // Loads are never behind a branch: an item past n_pairs (the last
    // blocks' tail) reads pair n_pairs - 1, the last item's prefetch reads
    // its own pair again, and ok only guards the stores.  A load behind a
    // branch would make the compiler wait for every load outstanding -- the
    // next pair's prefetch included -- before each use of a row's values.
    auto clamp = [&](int64_t p) { return p < n_pairs ? p : n_pairs - 1; };
    TwCache<N, EPT> twc;
    twc.load(tw, t);
    float2 ya[EPT], yb[EPT], w0[EPT], w1[EPT];
    op.load_y(clamp(first), t, true, ya, yb);
    if constexpr (!FIRST) op.load_w(clamp(first), 0, t, true, w0);
    const uint32_t live = op.live_mask(t);
    twc.fence();
    for (int r = 0; r < run_count; ++r) {
        const int64_t it = first + (int64_t)r * Cfg::G;
        const bool ok = it < n_pairs;
        if constexpr (!FIRST) op.load_w(clamp(it), 1, t, true, w1);
        float2 z[EPT];
        op.pack(ya, yb, t, z);
        fft_n<N, EPT, true>(z, re, im, t, twc);
        float va[EPT], vb[EPT];
#pragma unroll
        for (int m = 0; m < EPT; ++m) {
            va[m] = fmaxf(z[m].x, 1e-13f);
            vb[m] = fmaxf(z[m].y, 1e-13f);
        }
        op.store_v(it, t, ok, va, vb);
        op.transmit_row(w0, va, t, re, im, twc);
        op.store_w(it, 0, t, ok, live, w0);
        const int64_t nx = clamp(r + 1 < run_count ? it + Cfg::G : it);
        op.load_y(nx, t, true, ya, yb);
        if constexpr (!FIRST) op.load_w(nx, 0, t, true, w0);
        op.transmit_row(w1, vb, t, re, im, twc);
        op.store_w(it, 1, t, ok, live, w1);
    }
}

// Last slice of the fused path: propagator and lens in one multiplier.
// H P / nPix with both masks (2/3 propagator mask and aperture).
struct LensPropMul {
    PropagatorMul p;
    LensMul l;
    static constexpr bool kTable = true;
    __device__ __forceinline__ void prepare(float2* px, int tid, int nthreads) const {
        p.prepare(px, tid, nthreads);
    }
    __device__ __forceinline__ float2 column(int j) const { return p.column(j); }
    __device__ __forceinline__ float2 apply(float2 w, int i, int j, const float2* px,
                                            float2 py) const {
        // l applies exp(-i chi) / nPix with the aperture; p.apply already
        // carries 1/nPix, so undo one of them
        const float2 wp = p.apply(w, i, j, px, py);
        const float2 wl = l(wp, i, j);
        const float n = (float)(p.n1 * p.n2);
        return make_float2(wl.x * n, wl.y * n);
    }
};

}  // namespace emsim
