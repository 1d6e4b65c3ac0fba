// C ABI of the learned coarse-slice correction (declarations:
// include/emsim_hip.h).  Host side: parameter layout, constant DFT bases,
// workspace carving, and the launch sequences of the forward pass, the
// backward pass and the Adam step (kernels in kernels_fno.hpp).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/emsim_hip.h"
#include "host_common.hpp"
#include "kernels_fno.hpp"

using namespace emsim;
using emsim_host::fail;

namespace {

constexpr double kTwoPi = 6.283185307179586476925286766559;

// emsim_hip_fno_force_generic: run every forward pass through k_gemm only
bool g_fno_force_generic = false;

uint16_t host_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int64_t align64(int64_t x) { return (x + 63) & ~int64_t(63); }
size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int elementwise_grid(int64_t n) {
    int64_t g = (n + 255) / 256;
    return (int)(g < 1 ? 1 : (g > 16384 ? 16384 : g));
}

int launch_gemm(const emsim_gemm_desc& d, hipStream_t s) {
    if (d.m <= 0 || d.n <= 0 || d.k <= 0 || d.batch <= 0 || !d.a || !d.b || !d.c)
        return fail(EMSIM_HIP_EINVAL, "gemm: empty problem or null operand");
    if (d.k2 < 0 || (d.k2 > 0 && (!d.a2 || !d.b2)))
        return fail(EMSIM_HIP_EINVAL, "gemm: second operand pair missing");
    if (d.split_k < 0 || (d.split_k > 0 && d.c_type != EMSIM_DT_F32))
        return fail(EMSIM_HIP_EINVAL, "gemm: split-k partials must be float32");
    const int64_t splits = d.split_k > 0 ? (d.k + d.split_k - 1) / d.split_k : 1;
    const int64_t gx = (d.m + kGemmM - 1) / kGemmM, gy = (d.n + kGemmN - 1) / kGemmN;
    const int64_t gz = d.batch * splits;
    if (gx > 0x7fffffff || gy > 65535 || gz > 65535)
        return fail(EMSIM_HIP_EUNSUPPORTED, "gemm: grid %lld x %lld x %lld too large",
                    (long long)gx, (long long)gy, (long long)gz);
    EMSIM_LAUNCH(0, k_gemm, dim3((unsigned)gx, (unsigned)gy, (unsigned)gz), dim3(kGemmThreads), 0,
                 s, d);
    return EMSIM_HIP_OK;
}

emsim_gemm_desc gemm_desc() {
    emsim_gemm_desc d;
    std::memset(&d, 0, sizeof(d));
    d.batch = 1;
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------
// the network
// ---------------------------------------------------------------------------
struct emsim_hip_fno {
    int n, cin, cin_pad, C, modes, L, M1, M2;
    int64_t n_params;
    int64_t off_lw, off_lb, off_qw, off_qb;
    std::vector<int64_t> off_w, off_b, off_r;
    float* params = nullptr;      // float32 master copy
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    uint16_t* params16 = nullptr; // bf16 mirror (GEMM operands)
    uint16_t* fy = nullptr;       // (2 M2, n)
    uint16_t* fx2 = nullptr;      // (2 M1, n)
    uint16_t* gx = nullptr;       // (2 n, 2 M1)
    uint16_t* gy = nullptr;       // (n, 2 M2)
};

namespace {

struct FnoWs {
    uint16_t* v[9];      // activations (bf16, P x C): v[0] lifting .. v[L]
    uint16_t* z[8];      // pre-activations of each layer (training)
    float* ty;           // (B n, 2 M2 C): float32, or bf16 written by the fused rows
    float* pq[8];        // (B, 2 M1, 2 M2 C) per layer (training keeps all)
    float* ob;           // (B, 2 M1, M2 C)
    float* u;            // (B, 2 n, M2 C); bf16 in the fused forward
    // backward
    float* dv;           // (P, C)
    float* dz;           // (P, C)
    float* dout_partial; // split-k partials
    float* colsum_part;
    size_t bytes;
};

constexpr int kSplits = 512;       // split-k partial count for weight gradients
constexpr int kColsumBlocks = 256;

FnoWs carve(const emsim_hip_fno* f, int B, int training, char* base) {
    FnoWs w;
    std::memset(&w, 0, sizeof(w));
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + off : nullptr;
        off += align256(bytes);
        return p;
    };
    const int64_t P = (int64_t)B * f->n * f->n;
    const int nv = training ? f->L + 1 : 2;
    for (int i = 0; i < nv; ++i) w.v[i] = (uint16_t*)take((size_t)P * f->C * 2);
    if (training)
        for (int i = 0; i < f->L; ++i) w.z[i] = (uint16_t*)take((size_t)P * f->C * 2);
    const size_t ty = (size_t)B * f->n * 2 * f->M2 * f->C * 4;
    const size_t pq = (size_t)B * 2 * f->M1 * 2 * f->M2 * f->C * 4;
    const size_t ob = (size_t)B * 2 * f->M1 * f->M2 * f->C * 4;
    const size_t u = (size_t)B * 2 * f->n * f->M2 * f->C * 4;
    w.ty = (float*)take(ty);
    for (int i = 0; i < (training ? f->L : 1); ++i) w.pq[i] = (float*)take(pq);
    w.ob = (float*)take(ob);
    w.u = (float*)take(u);
    if (training) {
        w.dv = (float*)take((size_t)P * f->C * 4);
        w.dz = (float*)take((size_t)P * f->C * 4);
        const int64_t wmax = (int64_t)(f->cin_pad > f->C ? f->cin_pad : f->C) * f->C;
        w.dout_partial = (float*)take((size_t)kSplits * wmax * 4);
        w.colsum_part = (float*)take((size_t)kColsumBlocks * f->C * 4);
    }
    w.bytes = off;
    return w;
}

int colsum(const float* x, int64_t rows, int cols, float* part, float* out, hipStream_t s) {
    EMSIM_LAUNCH(0, k_colsum_partial, dim3(kColsumBlocks), dim3(256), 0, s, x, rows, cols, part);
    EMSIM_LAUNCH(0, k_rowsum, dim3(1), dim3(256), 0, s, (const float*)part, kColsumBlocks,
                 (int64_t)cols, out, 0);
    return EMSIM_HIP_OK;
}

// out (m x n) = sum over P rows of a(row, i) b(row, j): A^T B with split-k
int weight_grad(const void* a, int a_type, int64_t lda, const void* b, int b_type, int64_t ldb,
                int64_t P, int m, int n, float* partial, float* out, hipStream_t s) {
    emsim_gemm_desc d = gemm_desc();
    d.m = m;
    d.n = n;
    d.k = P;
    d.a = a;
    d.a_type = a_type;
    d.a_sm = 1;
    d.a_sk = lda;
    d.b = b;
    d.b_type = b_type;
    d.b_sk = ldb;
    d.b_sn = 1;
    d.c = partial;
    d.c_type = EMSIM_DT_F32;
    d.c_sm = n;
    d.c_sn = 1;
    d.split_k = ((P + kSplits - 1) / kSplits + kGemmK - 1) / kGemmK * kGemmK;
    d.c_split_stride = (int64_t)m * n;
    const int64_t splits = (P + d.split_k - 1) / d.split_k;
    int rc = launch_gemm(d, s);
    if (rc) return rc;
    EMSIM_LAUNCH(0, k_rowsum, dim3(elementwise_grid((int64_t)m * n)), dim3(256), 0, s,
                 (const float*)partial, (int)splits, (int64_t)m * n, out, 0);
    return EMSIM_HIP_OK;
}

}  // namespace

extern "C" {

int emsim_hip_gemm(const emsim_gemm_desc* d, emsim_stream_t stream) {
    if (!d) return fail(EMSIM_HIP_EINVAL, "gemm: null descriptor");
    return launch_gemm(*d, (hipStream_t)stream);
}

int emsim_hip_fno_create(emsim_hip_fno** out, int n, int in_channels, int channels, int modes,
                         int layers) {
    if (!out) return fail(EMSIM_HIP_EINVAL, "fno_create: null output");
    *out = nullptr;
    if (n < 16 || n > 4096 || (n & (n - 1)))
        return fail(EMSIM_HIP_EUNSUPPORTED, "fno_create: n = %d is not a power of two in [16, 4096]", n);
    if (channels != 16 && channels != 32 && channels != 64)
        return fail(EMSIM_HIP_EUNSUPPORTED, "fno_create: channels must be 16, 32 or 64");
    if (modes < 1 || 2 * modes > n / 2)
        return fail(EMSIM_HIP_EINVAL, "fno_create: modes = %d must be in [1, n / 4]", modes);
    if (layers < 1 || layers > 8) return fail(EMSIM_HIP_EINVAL, "fno_create: layers must be 1..8");
    if (in_channels < 1 || in_channels > 64)
        return fail(EMSIM_HIP_EINVAL, "fno_create: in_channels must be 1..64");
    emsim_hip_fno* f = new (std::nothrow) emsim_hip_fno();
    if (!f) return fail(EMSIM_HIP_ENOMEM, "fno_create: out of host memory");
    f->n = n;
    f->cin = in_channels;
    f->cin_pad = (in_channels + 15) / 16 * 16;
    f->C = channels;
    f->modes = modes;
    f->L = layers;
    f->M1 = 2 * modes;
    f->M2 = modes;
    int64_t off = 0;
    auto take = [&](int64_t count) {
        const int64_t o = off;
        off += align64(count);
        return o;
    };
    f->off_lw = take((int64_t)f->cin_pad * f->C);
    f->off_lb = take(f->C);
    for (int l = 0; l < layers; ++l) {
        f->off_w.push_back(take((int64_t)f->C * f->C));
        f->off_b.push_back(take(f->C));
        f->off_r.push_back(take((int64_t)f->M1 * f->M2 * f->C * f->C * 2));
    }
    f->off_qw = take((int64_t)f->C * 2);
    f->off_qb = take(2);
    f->n_params = off;

    auto fail_free = [&](int code, const char* what) {
        emsim_hip_fno_destroy(f);
        return fail(code, "fno_create: %s", what);
    };
    const size_t pb = (size_t)f->n_params * 4;
    if (hipMalloc(&f->params, pb) != hipSuccess || hipMalloc(&f->grads, pb) != hipSuccess ||
        hipMalloc(&f->adam_m, pb) != hipSuccess || hipMalloc(&f->adam_v, pb) != hipSuccess ||
        hipMalloc(&f->params16, (size_t)f->n_params * 2) != hipSuccess)
        return fail_free(EMSIM_HIP_ENOMEM, "device allocation failed");
    if (hipMemset(f->params, 0, pb) != hipSuccess || hipMemset(f->grads, 0, pb) != hipSuccess ||
        hipMemset(f->adam_m, 0, pb) != hipSuccess || hipMemset(f->adam_v, 0, pb) != hipSuccess ||
        hipMemset(f->params16, 0, (size_t)f->n_params * 2) != hipSuccess)
        return fail_free(EMSIM_HIP_EDEVICE, "device memset failed");

    // constant bases, float64 on the host, rounded once to bf16
    const int M1 = f->M1, M2 = f->M2;
    auto kv = [&](int j) { return j < modes ? j : n - M1 + j; };   // retained k1 values
    std::vector<uint16_t> fy((size_t)2 * M2 * n), fx2((size_t)2 * M1 * n), gx((size_t)2 * n * 2 * M1),
        gy((size_t)n * 2 * M2);
    const double inv_n2 = 1.0 / ((double)n * (double)n);
    for (int k2 = 0; k2 < M2; ++k2)
        for (int y = 0; y < n; ++y) {
            const double ph = kTwoPi * (double)((int64_t)k2 * y % n) / n;
            fy[(size_t)k2 * n + y] = host_bf16((float)std::cos(ph));
            fy[(size_t)(M2 + k2) * n + y] = host_bf16((float)-std::sin(ph));
            const double c = (k2 == 0 ? 1.0 : 2.0) * inv_n2;
            gy[(size_t)y * 2 * M2 + k2] = host_bf16((float)(c * std::cos(ph)));
            gy[(size_t)y * 2 * M2 + M2 + k2] = host_bf16((float)(-c * std::sin(ph)));
        }
    for (int j = 0; j < M1; ++j)
        for (int x = 0; x < n; ++x) {
            const double th = kTwoPi * (double)((int64_t)kv(j) * x % n) / n;
            const float c = (float)std::cos(th), s = (float)std::sin(th);
            fx2[(size_t)j * n + x] = host_bf16(c);
            fx2[(size_t)(M1 + j) * n + x] = host_bf16(s);
            gx[(size_t)(2 * x) * 2 * M1 + j] = host_bf16(c);
            gx[(size_t)(2 * x) * 2 * M1 + M1 + j] = host_bf16(-s);
            gx[(size_t)(2 * x + 1) * 2 * M1 + j] = host_bf16(s);
            gx[(size_t)(2 * x + 1) * 2 * M1 + M1 + j] = host_bf16(c);
        }
    auto upload = [&](uint16_t** dst, const std::vector<uint16_t>& src) {
        if (hipMalloc(dst, src.size() * 2) != hipSuccess) return false;
        return hipMemcpy(*dst, src.data(), src.size() * 2, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!upload(&f->fy, fy) || !upload(&f->fx2, fx2) || !upload(&f->gx, gx) || !upload(&f->gy, gy))
        return fail_free(EMSIM_HIP_EDEVICE, "basis upload failed");
    *out = f;
    return EMSIM_HIP_OK;
}

void emsim_hip_fno_destroy(emsim_hip_fno* f) {
    if (!f) return;
    for (void* p : {(void*)f->params, (void*)f->grads, (void*)f->adam_m, (void*)f->adam_v,
                    (void*)f->params16, (void*)f->fy, (void*)f->fx2, (void*)f->gx, (void*)f->gy})
        if (p) (void)hipFree(p);
    delete f;
}

int64_t emsim_hip_fno_param_count(const emsim_hip_fno* f) { return f ? f->n_params : -1; }

int emsim_hip_fno_param_offsets(const emsim_hip_fno* f, int64_t* offsets_host, int cap) {
    // lift_w, lift_b, {w_l, b_l, r_l}, proj_w, proj_b
    if (!f || !offsets_host) return fail(EMSIM_HIP_EINVAL, "fno_param_offsets: null argument");
    const int need = 4 + 3 * f->L;
    if (cap < need) return fail(EMSIM_HIP_EINVAL, "fno_param_offsets: need %d entries", need);
    int i = 0;
    offsets_host[i++] = f->off_lw;
    offsets_host[i++] = f->off_lb;
    for (int l = 0; l < f->L; ++l) {
        offsets_host[i++] = f->off_w[l];
        offsets_host[i++] = f->off_b[l];
        offsets_host[i++] = f->off_r[l];
    }
    offsets_host[i++] = f->off_qw;
    offsets_host[i++] = f->off_qb;
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_set_params(emsim_hip_fno* f, const float* params_dev, emsim_stream_t stream) {
    if (!f || !params_dev) return fail(EMSIM_HIP_EINVAL, "fno_set_params: null argument");
    hipStream_t s = (hipStream_t)stream;
    EMSIM_HIP_CHECK(hipMemcpyAsync(f->params, params_dev, (size_t)f->n_params * 4,
                                   hipMemcpyDeviceToDevice, s));
    EMSIM_LAUNCH(0, k_to_bf16, dim3(elementwise_grid(f->n_params)), dim3(256), 0, s,
                 (const float*)f->params, f->params16, f->n_params);
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_get_params(const emsim_hip_fno* f, float* params_dev, emsim_stream_t stream) {
    if (!f || !params_dev) return fail(EMSIM_HIP_EINVAL, "fno_get_params: null argument");
    EMSIM_HIP_CHECK(hipMemcpyAsync(params_dev, f->params, (size_t)f->n_params * 4,
                                   hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_get_grads(const emsim_hip_fno* f, float* grads_dev, emsim_stream_t stream) {
    if (!f || !grads_dev) return fail(EMSIM_HIP_EINVAL, "fno_get_grads: null argument");
    EMSIM_HIP_CHECK(hipMemcpyAsync(grads_dev, f->grads, (size_t)f->n_params * 4,
                                   hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return EMSIM_HIP_OK;
}

size_t emsim_hip_fno_workspace_size(const emsim_hip_fno* f, int batch, int training) {
    if (!f || batch < 1) return 0;
    return carve(f, batch, training, nullptr).bytes;
}

int emsim_hip_fno_pack_input(const emsim_hip_fno* f, const float* wave_dev,
                             const float* slices_dev, int n_slices, float psi0, float sigma,
                             int batch, void* x_dev, emsim_stream_t stream) {
    if (!f || !wave_dev || !x_dev || (n_slices > 0 && !slices_dev))
        return fail(EMSIM_HIP_EINVAL, "fno_pack_input: null argument");
    if (n_slices < 0 || 2 + n_slices != f->cin)
        return fail(EMSIM_HIP_EINVAL, "fno_pack_input: 2 + n_slices = %d != in_channels %d",
                    2 + n_slices, f->cin);
    if (batch < 1 || !(psi0 > 0.f)) return fail(EMSIM_HIP_EINVAL, "fno_pack_input: bad batch/psi0");
    const int64_t npix = (int64_t)f->n * f->n, total = npix * batch;
    EMSIM_LAUNCH(0, k_fno_pack, dim3(elementwise_grid(total)), dim3(256), 0, (hipStream_t)stream,
                 (const float2*)wave_dev, slices_dev, n_slices, npix, total, 1.f / psi0, sigma,
                 (uint16_t*)x_dev, f->cin_pad);
    return EMSIM_HIP_OK;
}

}  // extern "C"

namespace {

// The 32-channel network with at most 16 y-modes on rows of at most 512
// pixels runs its row-local work in k_fno_rows; anything else in k_gemm.
// columns per lane load of k_fno_xfwd (1, 2 or 4)
#ifndef EMSIM_XFWD_V
#define EMSIM_XFWD_V 2
#endif
constexpr int kXfwdV = EMSIM_XFWD_V;

bool fused_rows_ok(const emsim_hip_fno* f) {
    return f->C == 32 && f->n <= 512 && f->n % 32 == 0 && (2 * f->M2 == 16 || 2 * f->M2 == 32) &&
           f->cin_pad <= 32;
}

// x-transform, channel mix and inverse x-transform of layer l (T -> U).
// u16: the fused row kernels' bf16 T and U (they are rounded to bf16 where
// they are read in any case); then the x-transforms run in k_fno_xfwd /
// k_fno_xinv where the shapes allow, else in k_gemm.
int spectral_middle(const emsim_hip_fno* f, const FnoWs& w, int batch, int l, float* pq,
                    hipStream_t s, bool u16 = false) {
    const int n = f->n, C = f->C, M1 = f->M1, M2 = f->M2;
    int rc;
    // the fused path's own x-transform kernels where their shapes allow
    const bool xk = u16 && 2 * M1 <= 64 && 2 * M1 % 16 == 0 && n % 64 == 0 &&
                    (M2 * C == 256 || M2 * C == 512) && (int64_t)batch * 2 * n / 32 <= 0x7fffffff;
    if (xk) {
        const int nc = 2 * M2 * C;
        const dim3 grid((unsigned)((int64_t)batch * nc / (32 * kXfwdV)));
        if (2 * M1 <= 32)
            EMSIM_LAUNCH(0, (k_fno_xfwd<1, kXfwdV>), grid, dim3(256), 0, s, f->fx2,
                         (const uint16_t*)w.ty, pq, n, 2 * M1, nc);
        else
            EMSIM_LAUNCH(0, (k_fno_xfwd<2, kXfwdV>), grid, dim3(256), 0, s, f->fx2,
                         (const uint16_t*)w.ty, pq, n, 2 * M1, nc);
    } else {
        emsim_gemm_desc d = gemm_desc();
        d.m = 2 * M1; d.n = 2 * M2 * C; d.k = n; d.batch = batch;
        d.a = f->fx2; d.a_type = EMSIM_DT_BF16; d.a_sm = n; d.a_sk = 1;
        d.b = w.ty; d.b_type = u16 ? EMSIM_DT_BF16 : EMSIM_DT_F32;
        d.b_bs = (int64_t)n * 2 * M2 * C; d.b_sk = 2 * M2 * C; d.b_sn = 1;
        d.c = pq; d.c_type = EMSIM_DT_F32; d.c_bs = (int64_t)2 * M1 * 2 * M2 * C; d.c_sm = 2 * M2 * C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
    }
    const float2* R = (const float2*)(f->params + f->off_r[l]);
    if (C == 32)
        EMSIM_LAUNCH(0, k_fno_mix<32>, dim3(M1 * M2), dim3(256), 0, s, (const float*)pq, R, w.ob, batch, M1, M2);
    else if (C == 16)
        EMSIM_LAUNCH(0, k_fno_mix<16>, dim3(M1 * M2), dim3(256), 0, s, (const float*)pq, R, w.ob, batch, M1, M2);
    else
        EMSIM_LAUNCH(0, k_fno_mix<64>, dim3(M1 * M2), dim3(256), 0, s, (const float*)pq, R, w.ob, batch, M1, M2);
    if (xk) {
        const dim3 grid((unsigned)((int64_t)batch * 2 * n / 32));
        if (M2 * C == 512)
            EMSIM_LAUNCH(0, k_fno_xinv<512>, grid, dim3(256), 0, s, f->gx, (const float*)w.ob,
                         (uint16_t*)w.u, 2 * n, 2 * M1);
        else
            EMSIM_LAUNCH(0, k_fno_xinv<256>, grid, dim3(256), 0, s, f->gx, (const float*)w.ob,
                         (uint16_t*)w.u, 2 * n, 2 * M1);
        return EMSIM_HIP_OK;
    }
    emsim_gemm_desc d = gemm_desc();
    d.m = 2 * n; d.n = M2 * C; d.k = 2 * M1; d.batch = batch;
    d.a = f->gx; d.a_type = EMSIM_DT_BF16; d.a_sm = 2 * M1; d.a_sk = 1;
    d.b = w.ob; d.b_type = EMSIM_DT_F32; d.b_bs = (int64_t)2 * M1 * M2 * C; d.b_sk = M2 * C; d.b_sn = 1;
    d.c = w.u; d.c_type = u16 ? EMSIM_DT_BF16 : EMSIM_DT_F32; d.c_bs = (int64_t)2 * n * M2 * C; d.c_sm = M2 * C; d.c_sn = 1;
    return launch_gemm(d, s);
}

// the row-kernel modes the forward pass uses
constexpr int kModeLift = kRowV | kRowT;
constexpr int kModeLiftPacked = kModeLift | kRowPack;
constexpr int kModeMid = kRowU | kRowAct | kRowV | kRowT;
constexpr int kModeMidTrain = kModeMid | kRowZ;
constexpr int kModeLast = kRowU | kRowProj;
constexpr int kModeLastTrain = kModeLast | kRowV | kRowZ;
// inference from the simulation's outputs: the lifting only feeds the first
// layer's spectral branch (its activations are not written), and the first
// layer's row kernel recomputes them from the same inputs
constexpr int kModeLiftPackedT = kRowT | kRowPack;
constexpr int kModeMidRelift = kModeMid | kRowRelift;

template <int K1F, int K2F>
int launch_rows_mode(const FnoRowArgs& a, int mode, dim3 grid, size_t lds, hipStream_t s) {
    const dim3 block(256);
    switch (mode) {
    case kModeLift: EMSIM_LAUNCH(0, (k_fno_rows<K1F, K2F, kModeLift>), grid, block, lds, s, a); break;
    case kModeMid: EMSIM_LAUNCH(0, (k_fno_rows<K1F, K2F, kModeMid>), grid, block, lds, s, a); break;
    case kModeMidTrain: EMSIM_LAUNCH(0, (k_fno_rows<K1F, K2F, kModeMidTrain>), grid, block, lds, s, a); break;
    case kModeLast: EMSIM_LAUNCH(0, (k_fno_rows<K1F, K2F, kModeLast>), grid, block, lds, s, a); break;
    case kModeLastTrain: EMSIM_LAUNCH(0, (k_fno_rows<K1F, K2F, kModeLastTrain>), grid, block, lds, s, a); break;
    default: return fail(EMSIM_HIP_EINVAL, "fno rows: unsupported mode %d", mode);
    }
    return EMSIM_HIP_OK;
}

// Rows per row-kernel block: the Gy / Fy fragments and W a block loads serve
// this many rows.
#ifndef EMSIM_FNO_RPB
#define EMSIM_FNO_RPB 16
#endif

int launch_rows(const FnoRowArgs& a_in, int mode, int64_t rows, hipStream_t s) {
    FnoRowArgs a = a_in;
    a.rows = rows;
    a.rpb = EMSIM_FNO_RPB;
    const dim3 grid((unsigned)((rows + a.rpb - 1) / a.rpb));
    const size_t lds = kFnoRowLds(a.n);
    if (mode == kModeLiftPackedT) {
        if (a.k1 != 16) return fail(EMSIM_HIP_EINVAL, "fno rows: re-lifting takes 16 input channels");
        if (a.m2x2 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<1, 1, kModeLiftPackedT>), grid, dim3(256), lds, s, a);
        else
            EMSIM_LAUNCH(0, (k_fno_rows<1, 2, kModeLiftPackedT>), grid, dim3(256), lds, s, a);
        return EMSIM_HIP_OK;
    }
    if (mode == kModeMidRelift) {
        if (a.k1 != 32) return fail(EMSIM_HIP_EINVAL, "fno rows: re-lifting feeds 32 channels");
        const size_t lds_r = lds + kFnoReliftLds;
        if (a.m2x2 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<2, 1, kModeMidRelift>), grid, dim3(256), lds_r, s, a);
        else
            EMSIM_LAUNCH(0, (k_fno_rows<2, 2, kModeMidRelift>), grid, dim3(256), lds_r, s, a);
        return EMSIM_HIP_OK;
    }
    if (mode == kModeLiftPacked) {
        if (a.k1 == 16 && a.m2x2 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<1, 1, kModeLiftPacked>), grid, dim3(256), lds, s, a);
        else if (a.k1 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<1, 2, kModeLiftPacked>), grid, dim3(256), lds, s, a);
        else if (a.m2x2 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<2, 1, kModeLiftPacked>), grid, dim3(256), lds, s, a);
        else
            EMSIM_LAUNCH(0, (k_fno_rows<2, 2, kModeLiftPacked>), grid, dim3(256), lds, s, a);
        return EMSIM_HIP_OK;
    }
    if (a.k1 == 16) {
        if (mode != kModeLift) return fail(EMSIM_HIP_EINVAL, "fno rows: 16 input channels only lift");
        if (a.m2x2 == 16)
            EMSIM_LAUNCH(0, (k_fno_rows<1, 1, kModeLift>), grid, dim3(256), lds, s, a);
        else
            EMSIM_LAUNCH(0, (k_fno_rows<1, 2, kModeLift>), grid, dim3(256), lds, s, a);
        return EMSIM_HIP_OK;
    }
    return a.m2x2 == 16 ? launch_rows_mode<2, 1>(a, mode, grid, lds, s)
                        : launch_rows_mode<2, 2>(a, mode, grid, lds, s);
}

// The lifting reads either the packed input x, or (x null) the exit waves
// and potentials directly.
struct PackedInput {
    const float2* wave;
    const float* slices;
    int ns;
    float inv_psi0, sigma;
};

int forward_fused(const emsim_hip_fno* f, const FnoWs& w, const void* x, float* out, int batch,
                  int training, hipStream_t s, const PackedInput* packed = nullptr) {
    const int64_t rows = (int64_t)batch * f->n;
    FnoRowArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n = f->n;
    a.gy = f->gy;
    a.m2x2 = 2 * f->M2;
    a.T = (uint16_t*)w.ty;
    // lifting, and the y-transform of its output
    a.vin = (const uint16_t*)x;
    a.k1 = f->cin_pad;
    a.w16 = f->params16 + f->off_lw;
    a.bias = f->params + f->off_lb;
    a.vout = w.v[0];
    a.fy = f->fy;
    if (packed) {
        a.wave = packed->wave;
        a.slices = packed->slices;
        a.ns = packed->ns;
        a.inv_psi0 = packed->inv_psi0;
        a.sigma = packed->sigma;
    }
    // inference from the simulation's outputs with a 16-channel input and a
    // first middle layer: the first layer recomputes the lifting instead of
    // the lifting writing it (about 1 GB less traffic per 64 images at 512^2)
    const bool relift = packed && !training && f->cin_pad == 16 && f->C == 32 && f->L >= 2;
    if (relift) {
        a.vout = nullptr;
        a.w16_lift = f->params16 + f->off_lw;
        a.bias_lift = f->params + f->off_lb;
    }
    int rc = launch_rows(a, relift ? kModeLiftPackedT : (packed ? kModeLiftPacked : kModeLift),
                         rows, s);
    if (rc) return rc;
    for (int l = 0; l < f->L; ++l) {
        const bool last = l + 1 == f->L;
        if ((rc = spectral_middle(f, w, batch, l, training ? w.pq[l] : w.pq[0], s, true)))
            return rc;
        a.vin = training ? w.v[l] : w.v[l & 1];
        a.k1 = f->C;
        a.w16 = f->params16 + f->off_w[l];
        a.bias = f->params + f->off_b[l];
        a.U = (const uint16_t*)w.u;
        a.act = last ? 0 : 1;
        a.vout = training ? w.v[l + 1] : (last ? nullptr : w.v[(l + 1) & 1]);
        a.zout = training ? w.z[l] : nullptr;
        a.fy = last ? nullptr : f->fy;
        a.q16 = last ? f->params16 + f->off_qw : nullptr;
        a.qb = f->params + f->off_qb;
        a.out = out;
        const int mode = last ? (training ? kModeLastTrain : kModeLast)
                              : (training ? kModeMidTrain
                                          : (relift && l == 0 ? kModeMidRelift : kModeMid));
        if ((rc = launch_rows(a, mode, rows, s))) return rc;
    }
    return EMSIM_HIP_OK;
}

}  // namespace

extern "C" {

int emsim_hip_fno_forward(emsim_hip_fno* f, const void* x_dev, float* out_dev, int batch,
                          int training, void* workspace_dev, size_t workspace_bytes,
                          emsim_stream_t stream) {
    if (!f || !x_dev || !out_dev || !workspace_dev)
        return fail(EMSIM_HIP_EINVAL, "fno_forward: null argument");
    if (batch < 1) return fail(EMSIM_HIP_EINVAL, "fno_forward: batch < 1");
    const FnoWs w = carve(f, batch, training, (char*)workspace_dev);
    if (workspace_bytes < w.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "fno_forward: workspace %zu < %zu", workspace_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    if (fused_rows_ok(f) && !g_fno_force_generic)
        return forward_fused(f, w, x_dev, out_dev, batch, training, s);
    const int n = f->n, C = f->C, M2 = f->M2;
    const int64_t P = (int64_t)batch * n * n;
    int rc;
    // lifting
    {
        emsim_gemm_desc d = gemm_desc();
        d.m = P; d.n = C; d.k = f->cin_pad;
        d.a = x_dev; d.a_type = EMSIM_DT_BF16; d.a_sm = f->cin_pad; d.a_sk = 1;
        d.b = f->params16 + f->off_lw; d.b_type = EMSIM_DT_BF16; d.b_sk = C; d.b_sn = 1;
        d.c = w.v[0]; d.c_type = EMSIM_DT_BF16; d.c_sm = C; d.c_sn = 1;
        d.bias = f->params + f->off_lb;
        if ((rc = launch_gemm(d, s))) return rc;
    }
    for (int l = 0; l < f->L; ++l) {
        const uint16_t* vin = training ? w.v[l] : w.v[l & 1];
        uint16_t* vout = training ? w.v[l + 1] : w.v[(l + 1) & 1];
        float* pq = training ? w.pq[l] : w.pq[0];
        // T = sum_y exp(-i k2 y) v           per image row
        emsim_gemm_desc d = gemm_desc();
        d.m = 2 * M2; d.n = C; d.k = n; d.batch = (int64_t)batch * n;
        d.a = f->fy; d.a_type = EMSIM_DT_BF16; d.a_sm = n; d.a_sk = 1;
        d.b = vin; d.b_type = EMSIM_DT_BF16; d.b_bs = (int64_t)n * C; d.b_sk = C; d.b_sn = 1;
        d.c = w.ty; d.c_type = EMSIM_DT_F32; d.c_bs = 2 * M2 * C; d.c_sm = C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
        // P, Q = sum_x (cos, sin)(k1 x) T; per-mode mix; U = sum_k1 exp(+i k1 x) O
        if ((rc = spectral_middle(f, w, batch, l, pq, s))) return rc;
        // v' = act(v W + b + irfft_y(U))    per image row
        d = gemm_desc();
        d.m = n; d.n = C; d.k = C; d.batch = (int64_t)batch * n;
        d.a = vin; d.a_type = EMSIM_DT_BF16; d.a_bs = (int64_t)n * C; d.a_sm = C; d.a_sk = 1;
        d.b = f->params16 + f->off_w[l]; d.b_type = EMSIM_DT_BF16; d.b_sk = C; d.b_sn = 1;
        d.k2 = 2 * M2;
        d.a2 = f->gy; d.a2_type = EMSIM_DT_BF16; d.a2_sm = 2 * M2; d.a2_sk = 1;
        d.b2 = w.u; d.b2_type = EMSIM_DT_F32; d.b2_bs = 2 * M2 * C; d.b2_sk = C; d.b2_sn = 1;
        d.c = vout; d.c_type = EMSIM_DT_BF16; d.c_bs = (int64_t)n * C; d.c_sm = C; d.c_sn = 1;
        d.bias = f->params + f->off_b[l];
        d.act = l + 1 < f->L ? 1 : 0;
        if (training) {
            d.z = w.z[l];
            d.z_type = EMSIM_DT_BF16;
        }
        if ((rc = launch_gemm(d, s))) return rc;
    }
    // projection
    {
        emsim_gemm_desc d = gemm_desc();
        d.m = P; d.n = 2; d.k = C;
        d.a = training ? w.v[f->L] : w.v[f->L & 1]; d.a_type = EMSIM_DT_BF16; d.a_sm = C; d.a_sk = 1;
        d.b = f->params16 + f->off_qw; d.b_type = EMSIM_DT_BF16; d.b_sk = 2; d.b_sn = 1;
        d.c = out_dev; d.c_type = EMSIM_DT_F32; d.c_sm = 2; d.c_sn = 1;
        d.bias = f->params + f->off_qb;
        if ((rc = launch_gemm(d, s))) return rc;
    }
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_backward(emsim_hip_fno* f, const void* x_dev, const float* dout_dev, int batch,
                           void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream) {
    if (!f || !x_dev || !dout_dev || !workspace_dev)
        return fail(EMSIM_HIP_EINVAL, "fno_backward: null argument");
    if (batch < 1) return fail(EMSIM_HIP_EINVAL, "fno_backward: batch < 1");
    const FnoWs w = carve(f, batch, 1, (char*)workspace_dev);
    if (workspace_bytes < w.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "fno_backward: workspace %zu < %zu", workspace_bytes, w.bytes);
    hipStream_t s = (hipStream_t)stream;
    const int n = f->n, C = f->C, M1 = f->M1, M2 = f->M2, L = f->L;
    const int64_t P = (int64_t)batch * n * n;
    float* g = f->grads;
    int rc;
    // projection: dQw = v_L^T dout, dqb = colsum(dout), dv = dout Qw^T
    if ((rc = weight_grad(w.v[L], EMSIM_DT_BF16, C, dout_dev, EMSIM_DT_F32, 2, P, C, 2,
                          w.dout_partial, g + f->off_qw, s)))
        return rc;
    if ((rc = colsum(dout_dev, P, 2, w.colsum_part, g + f->off_qb, s))) return rc;
    {
        emsim_gemm_desc d = gemm_desc();
        d.m = P; d.n = C; d.k = 2;
        d.a = dout_dev; d.a_type = EMSIM_DT_F32; d.a_sm = 2; d.a_sk = 1;
        d.b = f->params16 + f->off_qw; d.b_type = EMSIM_DT_BF16; d.b_sk = 1; d.b_sn = 2;
        d.c = w.dv; d.c_type = EMSIM_DT_F32; d.c_sm = C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
    }
    // the spectral-path buffers of the forward pass are reused for gradients
    float* dty = w.ty;
    float* dob = w.ob;
    float* du = w.u;
    for (int l = L - 1; l >= 0; --l) {
        const int act = l + 1 < L ? 1 : 0;
        EMSIM_LAUNCH(0, k_fno_act_bwd, dim3(elementwise_grid(P * C)), dim3(256), 0, s,
                     (const float*)w.dv, (const uint16_t*)w.z[l], w.dz, P * C, act);
        if ((rc = weight_grad(w.v[l], EMSIM_DT_BF16, C, w.dz, EMSIM_DT_F32, C, P, C, C,
                              w.dout_partial, g + f->off_w[l], s)))
            return rc;
        if ((rc = colsum(w.dz, P, C, w.colsum_part, g + f->off_b[l], s))) return rc;
        // dU = Gy^T dz                      per image row
        emsim_gemm_desc d = gemm_desc();
        d.m = 2 * M2; d.n = C; d.k = n; d.batch = (int64_t)batch * n;
        d.a = f->gy; d.a_type = EMSIM_DT_BF16; d.a_sm = 1; d.a_sk = 2 * M2;
        d.b = w.dz; d.b_type = EMSIM_DT_F32; d.b_bs = (int64_t)n * C; d.b_sk = C; d.b_sn = 1;
        d.c = du; d.c_type = EMSIM_DT_F32; d.c_bs = 2 * M2 * C; d.c_sm = C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
        // dOb = Gx^T dU                     per image
        d = gemm_desc();
        d.m = 2 * M1; d.n = M2 * C; d.k = 2 * n; d.batch = batch;
        d.a = f->gx; d.a_type = EMSIM_DT_BF16; d.a_sm = 1; d.a_sk = 2 * M1;
        d.b = du; d.b_type = EMSIM_DT_F32; d.b_bs = (int64_t)2 * n * M2 * C; d.b_sk = M2 * C; d.b_sn = 1;
        d.c = dob; d.c_type = EMSIM_DT_F32; d.c_bs = (int64_t)2 * M1 * M2 * C; d.c_sm = M2 * C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
        // mix backward: dR (into the gradient), dPQ (into the layer's PQ buffer after use)
        const float2* R = (const float2*)(f->params + f->off_r[l]);
        float2* dR = (float2*)(g + f->off_r[l]);
        float* dpq = w.pq[l];   // overwritten in place: each block reads its mode's PQ
                                // entries into LDS before writing that mode's dPQ
        if (C == 32)
            EMSIM_LAUNCH(0, k_fno_mix_bwd<32>, dim3(M1 * M2), dim3(256), 0, s, (const float*)w.pq[l], R, (const float*)dob, dR, dpq, batch, M1, M2);
        else if (C == 16)
            EMSIM_LAUNCH(0, k_fno_mix_bwd<16>, dim3(M1 * M2), dim3(256), 0, s, (const float*)w.pq[l], R, (const float*)dob, dR, dpq, batch, M1, M2);
        else
            EMSIM_LAUNCH(0, k_fno_mix_bwd<64>, dim3(M1 * M2), dim3(256), 0, s, (const float*)w.pq[l], R, (const float*)dob, dR, dpq, batch, M1, M2);
        // dT = Fx2^T dPQ                    per image
        d = gemm_desc();
        d.m = n; d.n = 2 * M2 * C; d.k = 2 * M1; d.batch = batch;
        d.a = f->fx2; d.a_type = EMSIM_DT_BF16; d.a_sm = 1; d.a_sk = n;
        d.b = dpq; d.b_type = EMSIM_DT_F32; d.b_bs = (int64_t)2 * M1 * 2 * M2 * C; d.b_sk = 2 * M2 * C; d.b_sn = 1;
        d.c = dty; d.c_type = EMSIM_DT_F32; d.c_bs = (int64_t)n * 2 * M2 * C; d.c_sm = 2 * M2 * C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
        // dv = dz W^T + Fy^T dT             per image row
        d = gemm_desc();
        d.m = n; d.n = C; d.k = C; d.batch = (int64_t)batch * n;
        d.a = w.dz; d.a_type = EMSIM_DT_F32; d.a_bs = (int64_t)n * C; d.a_sm = C; d.a_sk = 1;
        d.b = f->params16 + f->off_w[l]; d.b_type = EMSIM_DT_BF16; d.b_sk = 1; d.b_sn = C;
        d.k2 = 2 * M2;
        d.a2 = f->fy; d.a2_type = EMSIM_DT_BF16; d.a2_sm = 1; d.a2_sk = n;
        d.b2 = dty; d.b2_type = EMSIM_DT_F32; d.b2_bs = 2 * M2 * C; d.b2_sk = C; d.b2_sn = 1;
        d.c = w.dv; d.c_type = EMSIM_DT_F32; d.c_bs = (int64_t)n * C; d.c_sm = C; d.c_sn = 1;
        if ((rc = launch_gemm(d, s))) return rc;
    }
    // lifting: dLw = X^T dv0, dLb = colsum(dv0)
    if ((rc = weight_grad(x_dev, EMSIM_DT_BF16, f->cin_pad, w.dv, EMSIM_DT_F32, C, P, f->cin_pad,
                          C, w.dout_partial, g + f->off_lw, s)))
        return rc;
    return colsum(w.dv, P, C, w.colsum_part, g + f->off_lb, s);
}

int emsim_hip_fno_infer(emsim_hip_fno* f, const float* wave_dev, const float* slices_dev,
                        int n_slices, float psi0, float sigma, float* out_dev, int batch,
                        void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream) {
    if (!f || !wave_dev || !out_dev || !workspace_dev || (n_slices > 0 && !slices_dev))
        return fail(EMSIM_HIP_EINVAL, "fno_infer: null argument");
    if (batch < 1 || !(psi0 > 0.f) || n_slices < 0 || 2 + n_slices != f->cin)
        return fail(EMSIM_HIP_EINVAL, "fno_infer: bad batch, psi0 or slice count");
    if (!fused_rows_ok(f) || g_fno_force_generic)
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "fno_infer: this network shape needs fno_pack_input + fno_forward");
    const FnoWs w = carve(f, batch, 0, (char*)workspace_dev);
    if (workspace_bytes < w.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "fno_infer: workspace %zu < %zu", workspace_bytes, w.bytes);
    const PackedInput p{(const float2*)wave_dev, slices_dev, n_slices, 1.f / psi0, sigma};
    return forward_fused(f, w, nullptr, out_dev, batch, 0, (hipStream_t)stream, &p);
}

int emsim_hip_fno_force_generic(int on) {
    g_fno_force_generic = on != 0;
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_adam(emsim_hip_fno* f, float lr, float beta1, float beta2, float eps,
                       float weight_decay, int step, emsim_stream_t stream) {
    if (!f || step < 1) return fail(EMSIM_HIP_EINVAL, "fno_adam: null handle or step < 1");
    const float c1 = 1.f - std::pow(beta1, (float)step), c2 = 1.f - std::pow(beta2, (float)step);
    EMSIM_LAUNCH(0, k_adam, dim3(elementwise_grid(f->n_params)), dim3(256), 0, (hipStream_t)stream,
                 f->params, (const float*)f->grads, f->adam_m, f->adam_v, f->params16,
                 f->n_params, lr, beta1, beta2, eps, c1, c2, weight_decay);
    return EMSIM_HIP_OK;
}

int emsim_hip_fno_apply(const emsim_hip_fno* f, const float* coarse_wave_dev, const float* out_dev,
                        float psi0, float* wave_dev, int batch, emsim_stream_t stream) {
    if (!f || !coarse_wave_dev || !out_dev || !wave_dev || batch < 1)
        return fail(EMSIM_HIP_EINVAL, "fno_apply: bad argument");
    const int64_t total = (int64_t)batch * f->n * f->n;
    EMSIM_LAUNCH(0, k_fno_unpack, dim3(elementwise_grid(total)), dim3(256), 0, (hipStream_t)stream,
                 (const float2*)coarse_wave_dev, out_dev, psi0, (float2*)wave_dev, total);
    return EMSIM_HIP_OK;
}

size_t emsim_hip_fno_loss_workspace_size(int64_t npix, int batch) {
    (void)npix;
    return align256((size_t)batch * 64 * 3 * 4);
}

int emsim_hip_fno_loss(const float* pred_dev, const float* target_dev, int64_t npix, int batch,
                       float psi0, float* loss_dev, float* dout_dev, void* workspace_dev,
                       size_t workspace_bytes, emsim_stream_t stream) {
    if (!pred_dev || !target_dev || !loss_dev || !workspace_dev || npix < 1 || batch < 1 ||
        batch > 65535)
        return fail(EMSIM_HIP_EINVAL, "fno_loss: bad argument");
    if (workspace_bytes < emsim_hip_fno_loss_workspace_size(npix, batch))
        return fail(EMSIM_HIP_EWORKSPACE, "fno_loss: workspace too small");
    hipStream_t s = (hipStream_t)stream;
    const int nparts = 64;
    float* part = (float*)workspace_dev;
    EMSIM_LAUNCH(0, k_fno_loss_partial, dim3(nparts, batch), dim3(256), 0, s,
                 (const float2*)pred_dev, (const float2*)target_dev, npix, nparts, part);
    EMSIM_LAUNCH(0, k_fno_loss_finish, dim3((batch + 255) / 256), dim3(256), 0, s,
                 (const float*)part, nparts, batch, loss_dev);
    if (dout_dev) {
        const int64_t total = npix * batch;
        EMSIM_LAUNCH(0, k_fno_loss_grad, dim3(elementwise_grid(total)), dim3(256), 0, s,
                     (const float2*)pred_dev, (const float2*)target_dev, (const float*)loss_dev,
                     npix, total, 2.f * psi0 / (float)batch, dout_dev);
    }
    return EMSIM_HIP_OK;
}

}  // extern "C"
