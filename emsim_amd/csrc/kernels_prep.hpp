// Device-side preparation of a batch: one molecule in many orientations.
//
// Instance b of the batch is R_b (x - c0) for every atom x of the base
// molecule (already sorted by element), then translated so that its
// bounding box is centred on the origin, as atoms.centralize does on the
// host: r + ((-r_min) - (r_max - r_min) / 2), in float32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emsim {

// float <-> unsigned key whose unsigned order is the float order
__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void k_bbox_init(uint32_t* bbox, int batch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 6 * batch) bbox[i] = (i % 6) < 3 ? 0xffffffffu : 0u;
}

// grid (ceil(n / 256), B); bbox (B, 6) keys: min x3 (start 0xffffffff),
// max x3 (start 0)
__global__ __launch_bounds__(256) void k_orient(const float* __restrict__ base, int64_t n_atoms,
                                                const float* __restrict__ rot, float c0x,
                                                float c0y, float c0z, float* __restrict__ out,
                                                uint32_t* __restrict__ bbox) {
    const int b = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float* r = rot + 9 * b;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n_atoms) {
        const float x = base[3 * i] - c0x, y = base[3 * i + 1] - c0y, z = base[3 * i + 2] - c0z;
        float* o = out + ((int64_t)b * n_atoms + i) * 3;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float v = r[3 * d] * x + r[3 * d + 1] * y + r[3 * d + 2] * z;
            o[d] = v;
            mn[d] = v;
            mx[d] = v;
        }
    }
    __shared__ float red[6][256];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = mn[d];
        red[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicMin(&bbox[6 * b + threadIdx.x], float_key(red[threadIdx.x][0]));
    else if (threadIdx.x < 6) atomicMax(&bbox[6 * b + threadIdx.x], float_key(red[threadIdx.x][0]));
}

// Centring shift of axis d for a molecule with bounding-box keys bb (6):
// (-r_min) - (r_max - r_min) / 2 in float32, as atoms.centralize computes it.
__device__ __forceinline__ float centring_shift(const uint32_t* __restrict__ bb, int d) {
    const float mn = key_float(bb[d]), mx = key_float(bb[3 + d]);
    return __fsub_rn(-mn, __fsub_rn(mx, mn) * 0.5f);
}

// Bounding boxes of the molecules of a batch: grid (chunks, batch), block
// (chunk c, molecule b) reduces atoms [mol_off[b] + c CH, + CH) of molecule b
// (CH = 256 x kBoxPerThread) and folds them into bbox (batch, 6) keys with six
// atomics.  Atoms whose element index (elem, nullable) is n_elems or more
// are not part of the molecule and are left out.
constexpr int kBoxPerThread = 16;
__global__ __launch_bounds__(256) void k_mol_bbox(const float* __restrict__ coords,
                                                  const int64_t* __restrict__ mol_off,
                                                  const uint8_t* __restrict__ elem, int n_elems,
                                                  uint32_t* __restrict__ bbox) {
    const int b = blockIdx.y;
    const int64_t lo = mol_off[b], hi = mol_off[b + 1];
    const int64_t a0 = lo + (int64_t)blockIdx.x * (256 * kBoxPerThread);
    if (a0 >= hi) return;   // block-uniform: no barrier is skipped by part of a block
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < kBoxPerThread; ++k) {
        const int64_t a = a0 + (int64_t)k * 256 + threadIdx.x;
        if (a < hi && (!elem || elem[a] < n_elems)) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float v = coords[3 * a + d];
                mn[d] = fminf(mn[d], v);
                mx[d] = fmaxf(mx[d], v);
            }
        }
    }
    __shared__ float red[6][256];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        red[d][threadIdx.x] = mn[d];
        red[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                red[d][threadIdx.x] = fminf(red[d][threadIdx.x], red[d][threadIdx.x + s]);
                red[3 + d][threadIdx.x] = fmaxf(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 3) atomicMin(&bbox[6 * b + threadIdx.x], float_key(red[threadIdx.x][0]));
    else if (threadIdx.x < 6) atomicMax(&bbox[6 * b + threadIdx.x], float_key(red[threadIdx.x][0]));
}

__global__ void k_recenter(float* __restrict__ out, int64_t n_atoms, int batch,
                           const uint32_t* __restrict__ bbox) {
    const int64_t total = (int64_t)batch * n_atoms * 3;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
        const int64_t b = k / (n_atoms * 3);
        const int d = (int)(k % 3);
        const float mn = key_float(bbox[6 * b + d]), mx = key_float(bbox[6 * b + 3 + d]);
        const float shift = __fsub_rn(-mn, __fsub_rn(mx, mn) * 0.5f);
        out[k] = __fadd_rn(out[k], shift);
    }
}

}  // namespace emsim
