// C ABI of the gfx950 multislice backend (declarations: include/emsim_hip.h).
//
// Host side only: argument checks, float32 constants computed in the
// reference's operation order, constant-table uploads, and launches of the
// kernels in kernels_wave.hpp / kernels_slice.hpp on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/emsim_hip.h"
#include "host_common.hpp"
#include "fft_lds.hpp"
#include "kernels_slice.hpp"
#include "kernels_wave.hpp"
#include "kernels_fused.hpp"
#include "kernels_prep.hpp"
#include "kernels_dft.hpp"

#include <complex>

using namespace emsim;

namespace emsim_host {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

// --------------------------------------------------------------------------
// launches: error check plus optional per-kernel event timing
// (emsim_hip_timing_enable / emsim_hip_timing_collect).  Events are recorded
// on the launch stream, immediately before and after the kernel.
// --------------------------------------------------------------------------
struct TimingRecord {
    const char* name;
    int n;
    hipEvent_t start, stop;
};
std::mutex g_timing_mutex;
bool g_timing_on = false;
std::vector<TimingRecord> g_timing_records;
std::vector<hipEvent_t> g_event_pool;

hipEvent_t pooled_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

hipEvent_t timing_begin(hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    if (!g_timing_on) return nullptr;
    hipEvent_t e = pooled_event();
    if (e && hipEventRecord(e, s) != hipSuccess) {
        g_event_pool.push_back(e);
        return nullptr;
    }
    return e;
}

void timing_end(hipEvent_t start, const char* name, int n, hipStream_t s) {
    if (!start) return;
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    hipEvent_t e = pooled_event();
    if (!e || hipEventRecord(e, s) != hipSuccess) {
        g_event_pool.push_back(start);
        if (e) g_event_pool.push_back(e);
        return;
    }
    g_timing_records.push_back({name, n, start, e});
}

}  // namespace emsim_host

using namespace emsim_host;

namespace {

// The reference's PI literal (WavePropagator.cpp:10) as float32.
constexpr float kPi32 = 3.14159265358979324f;

bool pow2_supported(int n) { return n >= 8 && n <= kTwN && (n & (n - 1)) == 0; }
// Lengths the chirp-z path takes: any n with a power-of-two convolution
// length 2n - 1 <= M <= kTwN.
constexpr int kChirpMaxN = kTwN / 2;
bool chirp_supported(int n) { return n >= 8 && n <= kChirpMaxN; }
bool size_supported(int n) { return pow2_supported(n) || chirp_supported(n); }

// --------------------------------------------------------------------------
// chirp-z plans (kernels_dft.hpp): c_m = exp(-i pi m^2 / n) and the transform
// of the circular chirp b, both computed in float64 on the host
// --------------------------------------------------------------------------
struct ChirpPlan {
    int n = 0, M = 0;
    float2* c_fwd = nullptr;   // (n) exp(-i pi m^2 / n)
    float2* c_inv = nullptr;   // (n) conj
    float2* b_fwd = nullptr;   // (M) FFT_M of conj(c_fwd) placed at m and M - m
    float2* b_inv = nullptr;   // (M) the same for the inverse transform
};

// In-place iterative radix-2 FFT in float64, forward sign exp(-2 pi i j k / M).
void host_fft(std::vector<std::complex<double>>& a) {
    const size_t m = a.size();
    for (size_t i = 1, j = 0; i < m; ++i) {
        size_t bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t len = 2; len <= m; len <<= 1) {
        for (size_t i = 0; i < m; i += len) {
            for (size_t k = 0; k < len / 2; ++k) {
                const double ang = -two_pi * (double)k / (double)len;
                const std::complex<double> w(std::cos(ang), std::sin(ang));
                const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
        }
    }
}

void chirp_plan_free(ChirpPlan& p) {
    for (float2** q : {&p.c_fwd, &p.c_inv, &p.b_fwd, &p.b_inv}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
    }
}

int chirp_plan_create(int n, ChirpPlan& p) {
    int M = 1;
    while (M < 2 * n - 1) M <<= 1;
    if (M < 8) M = 8;
    if (M > kTwN) return fail(EMSIM_HIP_EUNSUPPORTED, "chirp-z length %d needs M = %d > %d", n, M, kTwN);
    p.n = n;
    p.M = M;
    const double pi = 3.14159265358979323846264338327950288;
    std::vector<std::complex<double>> c(n);
    for (int m = 0; m < n; ++m) {
        const long long q = ((long long)m * m) % (2LL * n);   // m^2 mod 2n: exact angle
        c[m] = std::polar(1.0, -pi * (double)q / (double)n);
    }
    std::vector<float2> hc(n), hci(n), hb(M), hbi(M);
    for (int m = 0; m < n; ++m) {
        hc[m] = make_float2((float)c[m].real(), (float)c[m].imag());
        hci[m] = make_float2((float)c[m].real(), (float)-c[m].imag());
    }
    for (int dir = 0; dir < 2; ++dir) {
        // the convolution kernel: conj of the transform's chirp, circular
        std::vector<std::complex<double>> b(M, 0.0);
        for (int m = 0; m < n; ++m) {
            const std::complex<double> v = dir == 0 ? std::conj(c[m]) : c[m];
            b[m] = v;
            if (m > 0) b[M - m] = v;
        }
        host_fft(b);
        std::vector<float2>& dst = dir == 0 ? hb : hbi;
        for (int k = 0; k < M; ++k) dst[k] = make_float2((float)b[k].real(), (float)b[k].imag());
    }
    hipError_t err = hipSuccess;
    const std::vector<float2>* src[4] = {&hc, &hci, &hb, &hbi};
    float2** dst[4] = {&p.c_fwd, &p.c_inv, &p.b_fwd, &p.b_inv};
    for (int i = 0; i < 4 && err == hipSuccess; ++i) {
        err = hipMalloc(dst[i], src[i]->size() * sizeof(float2));
        if (err == hipSuccess)
            err = hipMemcpy(*dst[i], src[i]->data(), src[i]->size() * sizeof(float2),
                            hipMemcpyHostToDevice);
    }
    if (err != hipSuccess) {
        chirp_plan_free(p);
        return fail(EMSIM_HIP_EDEVICE, "chirp plan: %s", hipGetErrorString(err));
    }
    return EMSIM_HIP_OK;
}

// --------------------------------------------------------------------------
// per-device twiddle table w[m] = exp(-2 pi i m / 4096), float64 -> float32
// --------------------------------------------------------------------------
std::mutex g_tw_mutex;
std::vector<float2*> g_twiddles;

int get_twiddles(const float2** out) {
    int dev = 0;
    EMSIM_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if ((int)g_twiddles.size() <= dev) g_twiddles.resize(dev + 1, nullptr);
    if (!g_twiddles[dev]) {
        std::vector<float2> host(kTwN);
        const double two_pi = 6.283185307179586476925286766559;
        for (int m = 0; m < kTwN; ++m) {
            const double a = -two_pi * (double)m / (double)kTwN;
            host[m] = make_float2((float)std::cos(a), (float)std::sin(a));
        }
        float2* d = nullptr;
        EMSIM_HIP_CHECK(hipMalloc(&d, kTwN * sizeof(float2)));
        hipError_t err = hipMemcpy(d, host.data(), kTwN * sizeof(float2), hipMemcpyHostToDevice);
        if (err != hipSuccess) {
            (void)hipFree(d);
            return fail(EMSIM_HIP_EDEVICE, "twiddle upload failed: %s", hipGetErrorString(err));
        }
        g_twiddles[dev] = d;
    }
    *out = g_twiddles[dev];
    return EMSIM_HIP_OK;
}

// --------------------------------------------------------------------------
// numpy.arange(start, stop, step) for float64, as the reference's binning
// calls it (slice_builder_base.py:183-186): length ceil((stop-start)/step),
// element 0 = start, element 1 = start + step, element i >= 2 =
// start + i * ((start + step) - start).
// --------------------------------------------------------------------------
void numpy_arange(double start, double stop, double step, std::vector<double>& out) {
    out.clear();
    const double len_f = std::ceil((stop - start) / step);
    if (!(len_f > 0)) return;
    const long len = (long)len_f;
    out.resize(len);
    out[0] = start;
    if (len == 1) return;
    const double next = start + step;
    out[1] = next;
    const double delta = next - start;
    for (long i = 2; i < len; ++i) out[i] = start + (double)i * delta;
}

void reference_edges(int n, double d, std::vector<double>& e64) {
    const double start = -d * (double)(n / 2);
    const double end = start + (double)n * d;
    numpy_arange(start, end + d, d, e64);
}

// Smallest float32 >= x.
float round_up_f32(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Row stride of the half-spectrum workspace: H = n2/2+1 rounded up to a
// multiple of 16 complex, so that column tiles start on 64-byte boundaries.
int half_ld(int H) { return (H + 15) & ~15; }

int grid_for(int64_t n, int block) {
    int64_t g = ceil_div(n, block);
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (int)g;
}

}  // namespace

// ==========================================================================
// handles
// ==========================================================================
struct emsim_hip_slice_builder {
    int n_elems = 0, n_slices = 0, n1 = 0, n2 = 0, H = 0;
    double dz = 0, d1 = 0, d2 = 0;
    int one_slice = 0;
    float* sf = nullptr;        // (n_elems, n1, H)
    // the same, tile-major for the fused spectrum pass: (n_elems, tiles, n1,
    // C) with C = spec_tile_cols(n1) columns per tile, zeros beyond H
    float* sf_tiles = nullptr;
    float* edges = nullptr;     // axis0 (n0+1) | axis1 (n1+1) | axis2 (n2+1)
    int last_exact[3] = {0, 0, 0};
    float inv_d[3] = {0, 0, 0};
    int tile_rows = 1, tiles_per_plane = 1;
    const float2* tw = nullptr;
    // a lateral size that is not a power of two: chirp-z plans of both axes
    bool chirp = false;
    ChirpPlan p1, p2;
};

namespace {
// Rows of the fused path's Y also hold the spectrum kernel's per-element
// column-0 transforms in padding columns H .. H + n_elems - 1.
int fused_ld(const emsim_hip_slice_builder* b) { return half_ld(b->H + b->n_elems); }
}  // namespace

struct emsim_hip_wave_propagator {
    int n1 = 0, n2 = 0;
    float d1 = 0, d2 = 0, lam = 0, gamma = 0;
    float dfx = 0, dfy = 0, filter2 = 0, factor = 0;
    const float2* tw = nullptr;
    // a wave shape that is not a power of two: chirp-z plans of both axes and
    // the line buffer of one image (max(n1 M2, n2 M1) complex64)
    bool chirp = false;
    ChirpPlan p1, p2;
    float2* work = nullptr;
};

// ==========================================================================
// size dispatch
// ==========================================================================
#define EMSIM_DISPATCH(n, FN, ...)                                                  \
    switch (n) {                                                                    \
        case 8: return FN<8>(__VA_ARGS__);                                          \
        case 16: return FN<16>(__VA_ARGS__);                                        \
        case 32: return FN<32>(__VA_ARGS__);                                        \
        case 64: return FN<64>(__VA_ARGS__);                                        \
        case 128: return FN<128>(__VA_ARGS__);                                      \
        case 256: return FN<256>(__VA_ARGS__);                                      \
        case 512: return FN<512>(__VA_ARGS__);                                      \
        case 1024: return FN<1024>(__VA_ARGS__);                                    \
        case 2048: return FN<2048>(__VA_ARGS__);                                    \
        case 4096: return FN<4096>(__VA_ARGS__);                                    \
        default: return fail(EMSIM_HIP_EUNSUPPORTED, "unsupported FFT length %d", n); \
    }

namespace {

// Launch-shape knobs (emsim_hip_tuning_set); 0 selects the built-in rule.
// They change how work is split over blocks, never the arithmetic.
int g_row_items_per_block = 0;   // row groups per row-worker block
int g_col_tiles_per_block = 0;   // column tiles per column-kernel block
int g_col_shape = 0;             // column-tile shape (ColCfg SHAPE) of the propagator pass

// Grid of the row worker for n_items rows (or row pairs) of length N.  With
// resident > 0 (blocks of this kernel the device holds at once), the fewest
// rows per block that still fit the grid in one round of resident blocks, so
// no partial second round trails the first; otherwise rows_per_block.
template <int N>
void row_grid(int64_t n_items, int& rpw, dim3& grid, int64_t resident = 0) {
    using Cfg = RowCfg<N>;
    const int64_t groups = ceil_div(n_items, Cfg::G);
    if (g_row_items_per_block > 0) {
        rpw = g_row_items_per_block;
    } else if (resident > 0) {
        const int64_t r = ceil_div(groups, resident);
        rpw = (int)(r < 1 ? 1 : (r > 16 ? 16 : r));
    } else {
        rpw = rows_per_block(groups);
    }
    grid = dim3((unsigned)ceil_div(groups, rpw));
}

// Blocks of kernel k (block size bs, no dynamic LDS) resident on the current
// device at once, from its occupancy and CU count; 0 if unknown.  Cached per
// device (a race between threads stores the same value).
template <class K>
int64_t resident_blocks(K k, int bs, int (&cache)[64]) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    if (cache[dev] == 0) {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, bs, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        cache[dev] = per_cu > 0 && cus > 0 ? per_cu * cus : -1;
    }
    return cache[dev] > 0 ? cache[dev] : 0;
}

template <int N>
int launch_row_fft(bool inv, const float2* in, float2* out, int64_t rows, const float2* tw,
                   hipStream_t s) {
    int rpw;
    dim3 grid;
    row_grid<N>(rows, rpw, grid);
    const dim3 block(RowCfg<N>::BLOCK);
    if (inv) EMSIM_LAUNCH(N, (k_row_fft<N, true>), grid, block, 0, s, OpFFT<N, true>{in, out}, rows, rpw, tw);
    else EMSIM_LAUNCH(N, (k_row_fft<N, false>), grid, block, 0, s, OpFFT<N, false>{in, out}, rows, rpw, tw);
    return EMSIM_HIP_OK;
}

template <int N>
int launch_row_transmit_fft(bool first, const float2* in, const float* V, float2* out,
                            int64_t rows, int rows_per_img, int64_t v_img_stride, float factor,
                            const float2* tw, hipStream_t s) {
    static int resident_cache[64];
    const int64_t resident =
        resident_blocks(k_row_transmit_fft<N, false>, RowCfg<N>::BLOCK, resident_cache);
    int rpw;
    dim3 grid;
    row_grid<N>(rows, rpw, grid, resident);
    const dim3 block(RowCfg<N>::BLOCK);
    if (first)
        EMSIM_LAUNCH(N, (k_row_transmit_fft<N, true>), grid, block, 0, s,
                     (OpTransmit<N, true>{in, V, out, rows_per_img, v_img_stride, factor}), rows,
                     rpw, tw);
    else
        EMSIM_LAUNCH(N, (k_row_transmit_fft<N, false>), grid, block, 0, s,
                     (OpTransmit<N, false>{in, V, out, rows_per_img, v_img_stride, factor}), rows,
                     rpw, tw);
    return EMSIM_HIP_OK;
}

template <int N>
int launch_row_ifft_intensity(const float2* in, float* out, int64_t rows, const float2* tw,
                              hipStream_t s) {
    static int resident_cache[64];
    const int64_t resident =
        resident_blocks(k_row_ifft_intensity<N>, RowCfg<N>::BLOCK, resident_cache);
    int rpw;
    dim3 grid;
    row_grid<N>(rows, rpw, grid, resident);
    EMSIM_LAUNCH(N, (k_row_ifft_intensity<N>), grid, dim3(RowCfg<N>::BLOCK), 0, s,
                 (OpIfftIntensity<N>{in, out}), rows, rpw, tw);
    return EMSIM_HIP_OK;
}

// Wave columns [lo, hi] that the multiplier zeroes whole (none by default):
// pass B loads none of them, and with skip_tiles leaves out the column
// tiles that hold only such columns (their contents are then left as they
// are, for callers that never read them).
struct ColDead {
    int lo = 1 << 30, hi = -1;
    bool skip_tiles = false;
};

template <int N, class Mul, int SHAPE = 0>
int launch_col_mul_shape(float2* data, int ld, int ncols, int64_t img_stride, int n_img, Mul mul,
                         const float2* tw, hipStream_t s, ColDead cd) {
    using Cfg = ColCfg<N, SHAPE>;
    const int tiles_per_img = (int)ceil_div(ncols, Cfg::C);
    DeadCols dc;
    dc.lo = cd.lo;
    dc.hi = cd.hi;
    dc.live_per_img = tiles_per_img;
    dc.skip_lo = tiles_per_img;
    dc.skip_n = 0;
    if (cd.skip_tiles && cd.lo <= cd.hi) {
        // tiles [first, end) lie inside [lo, hi]
        const int first = (int)ceil_div(cd.lo, Cfg::C), end = (cd.hi + 1) / Cfg::C;
        if (end > first) {
            dc.skip_lo = first;
            dc.skip_n = end - first;
            dc.live_per_img = tiles_per_img - dc.skip_n;
        }
    }
    const int64_t n_tiles = (int64_t)dc.live_per_img * n_img;
    if (n_tiles <= 0) return EMSIM_HIP_OK;   // nothing to transform
    if (n_tiles >= ((int64_t)1 << 31))
        return fail(EMSIM_HIP_EINVAL, "column pass: %lld column tiles, at most 2^31 - 1",
                    (long long)n_tiles);
    // Blocks pair up on an XCD over strips of two tiles (see
    // k_col_fft_mul_ifft).  By default one round of resident blocks (two
    // per CU) with the strips dealt out evenly, or one strip per pair when
    // there are fewer strips than that; the tiles-per-block knob instead
    // sets the strips per pair.  Pairs come in eights (a grid of whole
    // multiples of 16).
    static int resident_cache[64];
    const int64_t strips = ceil_div(n_tiles, 2);
    int64_t pairs;
    if (g_col_tiles_per_block > 0) {
        pairs = ceil_div(strips, g_col_tiles_per_block);
    } else {
        const int64_t resident =
            resident_blocks(k_col_fft_mul_ifft<N, Mul, SHAPE>, Cfg::BLOCK, resident_cache);
        const int64_t max_pairs = resident >= 16 ? resident / 16 * 8 : 256;
        pairs = strips < max_pairs ? strips : max_pairs;
    }
    pairs = ceil_div(pairs, 8) * 8;
    const RunSched sched{(int)(strips / pairs), (int)(strips % pairs)};
    const dim3 grid((unsigned)(2 * pairs));
    EMSIM_LAUNCH(N, (k_col_fft_mul_ifft<N, Mul, SHAPE>), grid, dim3(Cfg::BLOCK), 0, s, data, ld,
                 ncols, dc, img_stride, n_tiles, sched, mul, tw);
    return EMSIM_HIP_OK;
}

// The propagator and last-slice passes of the fused path take the tuned tile
// shape (emsim_hip_tuning_set("col_shape")) at the lengths that have them.
template <int N, class Mul>
int launch_col_mul_impl(float2* data, int ld, int ncols, int64_t img_stride, int n_img, Mul mul,
                        const float2* tw, hipStream_t s, ColDead cd = ColDead{},
                        bool tuned = false) {
    if constexpr (col_shapes(N)) {
        if (tuned && g_col_shape == 1)
            return launch_col_mul_shape<N, Mul, 1>(data, ld, ncols, img_stride, n_img, mul, tw, s,
                                                   cd);
        if (tuned && g_col_shape == 2)
            return launch_col_mul_shape<N, Mul, 2>(data, ld, ncols, img_stride, n_img, mul, tw, s,
                                                   cd);
    }
    return launch_col_mul_shape<N, Mul, 0>(data, ld, ncols, img_stride, n_img, mul, tw, s, cd);
}

template <int N>
int launch_col_prop(float2* data, int ld, int ncols, int64_t img_stride, int n_img,
                    PropagatorMul mul, const float2* tw, hipStream_t s, ColDead cd) {
    return launch_col_mul_impl<N>(data, ld, ncols, img_stride, n_img, mul, tw, s, cd, true);
}

template <int N>
int launch_col_lens(float2* data, int ld, int ncols, int64_t img_stride, int n_img, LensMul mul,
                    const float2* tw, hipStream_t s) {
    return launch_col_mul_impl<N>(data, ld, ncols, img_stride, n_img, mul, tw, s);
}

template <int N>
int launch_r2c_pairs(const float* in, float2* out, int64_t n_pairs, int rows_per_plane,
                     int planes_per_group, int64_t in_group_stride, int64_t in_plane0, int ldh,
                     const float2* tw, hipStream_t s) {
    int rpw;
    dim3 grid;
    row_grid<N>(n_pairs, rpw, grid);
    OpR2CPairs<N> op;
    op.in = in;
    op.out = out;
    op.rows_per_plane = rows_per_plane;
    op.planes_per_group = planes_per_group;
    op.ldh = ldh;
    op.in_group_stride = in_group_stride;
    op.in_plane0 = in_plane0;
    EMSIM_LAUNCH(N, (k_row_r2c_pairs<N>), grid, dim3(RowCfg<N>::BLOCK), 0, s, op, n_pairs, rpw,
                 tw);
    return EMSIM_HIP_OK;
}

template <int N>
int launch_c2r_pairs(const float2* in, float* out, int64_t n_pairs, float clip_lo, int ldh,
                     const float2* tw, hipStream_t s) {
    int rpw;
    dim3 grid;
    row_grid<N>(n_pairs, rpw, grid);
    EMSIM_LAUNCH(N, (k_row_c2r_pairs<N>), grid, dim3(RowCfg<N>::BLOCK), 0, s,
                 (OpC2RPairs<N>{in, out, clip_lo, ldh}), n_pairs, rpw, tw);
    return EMSIM_HIP_OK;
}

template <int N>
int launch_col_spectrum(float2* X, const float* sf, int n_elems, int n_sc, int H, int ldh,
                        int n2, int bandlimit, int band_r2, float inv_npix, const float2* tw,
                        hipStream_t s) {
    using Cfg = SpecCfg<N>;
    const dim3 grid((unsigned)ceil_div(H, Cfg::C), (unsigned)n_sc);
    EMSIM_LAUNCH(N, (k_col_spectrum<N>), grid, dim3(Cfg::BLOCK), 0, s, X, sf, n_elems, n_sc, H,
                       ldh, n2, bandlimit, band_r2, inv_npix, tw);
    return EMSIM_HIP_OK;
}

int row_fft(int n, bool inv, const float2* in, float2* out, int64_t rows, const float2* tw,
            hipStream_t s) {
    EMSIM_DISPATCH(n, launch_row_fft, inv, in, out, rows, tw, s);
}
int row_transmit_fft(int n, bool first, const float2* in, const float* V, float2* out,
                     int64_t rows, int rows_per_img, int64_t v_img_stride, float factor,
                     const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_row_transmit_fft, first, in, V, out, rows, rows_per_img,
                   v_img_stride, factor, tw, s);
}
int row_ifft_intensity(int n, const float2* in, float* out, int64_t rows, const float2* tw,
                       hipStream_t s) {
    EMSIM_DISPATCH(n, launch_row_ifft_intensity, in, out, rows, tw, s);
}
int col_prop(int n, float2* data, int ld, int ncols, int64_t img_stride, int n_img,
             PropagatorMul mul, const float2* tw, hipStream_t s, ColDead cd = ColDead{}) {
    EMSIM_DISPATCH(n, launch_col_prop, data, ld, ncols, img_stride, n_img, mul, tw, s, cd);
}
int col_lens(int n, float2* data, int ld, int ncols, int64_t img_stride, int n_img, LensMul mul,
             const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_col_lens, data, ld, ncols, img_stride, n_img, mul, tw, s);
}
int r2c_pairs(int n, const float* in, float2* out, int64_t n_pairs, int rows_per_plane,
              int planes_per_group, int64_t in_group_stride, int64_t in_plane0, int ldh,
              const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_r2c_pairs, in, out, n_pairs, rows_per_plane, planes_per_group,
                   in_group_stride, in_plane0, ldh, tw, s);
}
int c2r_pairs(int n, const float2* in, float* out, int64_t n_pairs, float clip_lo, int ldh,
              const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_c2r_pairs, in, out, n_pairs, clip_lo, ldh, tw, s);
}
int col_spectrum(int n, float2* X, const float* sf, int n_elems, int n_sc, int H, int ldh,
                 int n2, int bandlimit, int band_r2, float inv_npix, const float2* tw,
                 hipStream_t s) {
    EMSIM_DISPATCH(n, launch_col_spectrum, X, sf, n_elems, n_sc, H, ldh, n2, bandlimit, band_r2,
                   inv_npix, tw, s);
}

// Float32 constants of the propagators, in the reference's operation order
// (WavePropagator.cpp:19-23,62-85,108-130).
PropagatorMul make_propagator(const emsim_hip_wave_propagator* w, float dz) {
    PropagatorMul p;
    p.n1 = w->n1;
    p.n2 = w->n2;
    p.dfx = w->dfx;
    p.dfy = w->dfy;
    p.filter2 = w->filter2;
    const float lp = w->lam * kPi32;
    p.lpd = lp * dz;
    p.inv_npix = 1.0f / (float)(w->n1 * w->n2);
    return p;
}

// The wave columns the propagator zeroes whole.  Its mask keeps
// f^2 = fx^2 + fy^2 <= filter^2 (PropagatorMul::apply), and in float32
// round-to-nearest f^2 >= fy^2, so every column with fy^2 > filter^2 is
// zero; fy^2 is formed by the same float32 operations as on the device.
// Reported only as one contiguous run of column indices (else none).
ColDead dead_columns(const PropagatorMul& p) {
    ColDead d;
    int lo = -1, hi = -1;
    for (int j = 0; j < p.n2; ++j) {
        const float fy = (float)freq_int(j, p.n2) * p.dfy;
        const float fy2 = fy * fy;
        if (fy2 <= p.filter2) continue;
        if (lo < 0) lo = j;
        else if (hi != j - 1) return ColDead{};
        hi = j;
    }
    if (lo >= 0) {
        d.lo = lo;
        d.hi = hi;
    }
    return d;
}

LensMul make_lens(const emsim_hip_wave_propagator* w, float cs_mm, float defocus, float aperture) {
    LensMul l;
    l.n1 = w->n1;
    l.n2 = w->n2;
    l.dfx = w->dfx;
    l.dfy = w->dfy;
    float c1 = 0.5f * kPi32;
    c1 = c1 * cs_mm;
    c1 = c1 * 1e7f;
    c1 = c1 * w->lam;
    c1 = c1 * w->lam;
    c1 = c1 * w->lam;
    float c2 = kPi32 * defocus;
    c2 = c2 * w->lam;
    const float f_aper = aperture / w->lam;
    l.c1 = c1;
    l.c2 = c2;
    l.faper2 = f_aper * f_aper;
    l.inv_npix = 1.0f / (float)(w->n1 * w->n2);
    return l;
}

// --------------------------------------------------------------------------
// chirp-z transforms of lines and images (sizes that are not powers of two)
// --------------------------------------------------------------------------
constexpr int64_t kOneImage = (int64_t)1 << 62;   // LineMap::per_img of a single image

// `lines` lines of p.n values (read from in through im in MODE, written to
// out through om): their unnormalised DFT (inv: exp(+2 pi i j k / n)),
// values k < n_out stored; real_out: real parts clipped at clip_lo.
// work: lines x p.M complex64.
int line_dft(const ChirpPlan& p, bool inv, int mode, const void* in, LineMap im, int64_t lines,
             void* out, LineMap om, int n_out, bool real_out, float clip_lo, float2* work,
             const float2* tw, hipStream_t s) {
    if (lines <= 0) return EMSIM_HIP_OK;
    const float2* c = inv ? p.c_inv : p.c_fwd;
    const float2* B = inv ? p.b_inv : p.b_fwd;
    const int64_t total = lines * (int64_t)p.M;
    const dim3 gi(grid_for(total, 256)), blk(256);
    if (mode == kReal)
        EMSIM_LAUNCH(p.n, k_chirp_in<kReal>, gi, blk, 0, s, in, im, lines, p.n, p.M, c, work);
    else if (mode == kHalf)
        EMSIM_LAUNCH(p.n, k_chirp_in<kHalf>, gi, blk, 0, s, in, im, lines, p.n, p.M, c, work);
    else
        EMSIM_LAUNCH(p.n, k_chirp_in<kComplex>, gi, blk, 0, s, in, im, lines, p.n, p.M, c, work);
    int rc = row_fft(p.M, false, work, work, lines, tw, s);
    if (rc) return rc;
    EMSIM_LAUNCH(p.n, k_line_mul, gi, blk, 0, s, work, lines, p.M, B);
    rc = row_fft(p.M, true, work, work, lines, tw, s);
    if (rc) return rc;
    const float scale = 1.0f / (float)p.M;   // exact: M is a power of two
    const dim3 go(grid_for(lines * (int64_t)n_out, 256));
    if (real_out)
        EMSIM_LAUNCH(p.n, k_chirp_out<true>, go, blk, 0, s, (const float2*)work, lines, n_out,
                     p.M, c, scale, clip_lo, out, om);
    else
        EMSIM_LAUNCH(p.n, k_chirp_out<false>, go, blk, 0, s, (const float2*)work, lines, n_out,
                     p.M, c, scale, clip_lo, out, om);
    return EMSIM_HIP_OK;
}

// Line maps of one n1 x n2 image with rows ld apart: its rows, its columns.
LineMap rows_of(int ld) { return LineMap{kOneImage, 0, ld, 1}; }
LineMap cols_of(int ld) { return LineMap{kOneImage, 0, 1, ld}; }

// Unnormalised 2-D DFT of one n1 x n2 complex image (out may be in).
int chirp_dft2(const emsim_hip_wave_propagator* w, bool inv, const float2* in, float2* out,
               hipStream_t s) {
    int rc = line_dft(w->p2, inv, kComplex, in, rows_of(w->n2), w->n1, out, rows_of(w->n2),
                      w->n2, false, 0.f, w->work, w->tw, s);
    if (rc) return rc;
    return line_dft(w->p1, inv, kComplex, out, cols_of(w->n2), w->n2, out, cols_of(w->n2),
                    w->n1, false, 0.f, w->work, w->tw, s);
}

// FFT -> Fourier-space multiplier -> IFFT of every image of a batch.
template <class Mul>
int chirp_filter(const emsim_hip_wave_propagator* w, const float2* in, float2* out, int batch,
                 const Mul& mul, hipStream_t s) {
    const int64_t npix = (int64_t)w->n1 * w->n2;
    for (int b = 0; b < batch; ++b) {
        float2* o = out + b * npix;
        int rc = chirp_dft2(w, false, in + b * npix, o, s);
        if (rc) return rc;
        EMSIM_LAUNCH(0, k_spectrum_mul<Mul>, dim3(grid_for(npix, 256)), dim3(256), 0, s, o, npix,
                     w->n1, w->n2, mul);
        rc = chirp_dft2(w, true, o, o, s);
        if (rc) return rc;
    }
    return EMSIM_HIP_OK;
}

}  // namespace

// ==========================================================================
// library
// ==========================================================================
extern "C" int emsim_hip_abi_version(void) { return EMSIM_HIP_ABI_VERSION; }

extern "C" const char* emsim_hip_last_error(void) { return g_last_error.c_str(); }

extern "C" int emsim_hip_fft_size_supported(int n) {
    return pow2_supported(n) ? 1 : (chirp_supported(n) ? 2 : 0);
}

extern "C" int emsim_hip_bin_edges(int n, double d, double* edges64_host, float* edges32_host,
                                   int cap, int* len_out) {
    if (n < 1 || !(d > 0) || !len_out) return fail(EMSIM_HIP_EINVAL, "bin_edges: bad arguments");
    std::vector<double> e;
    reference_edges(n, d, e);
    *len_out = (int)e.size();
    for (int i = 0; i < (int)e.size() && i < cap; ++i) {
        if (edges64_host) edges64_host[i] = e[i];
        if (edges32_host) edges32_host[i] = round_up_f32(e[i]);
    }
    return EMSIM_HIP_OK;
}

// ==========================================================================
// slice builder
// ==========================================================================
extern "C" int emsim_hip_slice_builder_create(emsim_hip_slice_builder** out,
                                              const float* sf_host, int n_elems, int n_slices,
                                              int n1, int n2, double dz, double d1, double d2,
                                              int one_slice) {
    if (!out) return fail(EMSIM_HIP_EINVAL, "slice_builder_create: null output handle");
    *out = nullptr;
    if (!sf_host || n_elems < 1 || n_slices < 1)
        return fail(EMSIM_HIP_EINVAL, "slice_builder_create: need n_elems >= 1, n_slices >= 1 "
                                      "and a scattering-factor table");
    if (n_elems > kMaxElems)
        return fail(EMSIM_HIP_EUNSUPPORTED, "slice_builder_create: at most %d elements", kMaxElems);
    if (!size_supported(n1) || !size_supported(n2))
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "slice_builder_create: lateral size %dx%d; the hip backend implements "
                    "sizes from 8 to %d (powers of two up to %d)", n1, n2, kChirpMaxN, kTwN);
    if (one_slice && n_slices != 1)
        return fail(EMSIM_HIP_EINVAL, "slice_builder_create: one-slice builder needs n_slices == 1");
    if (!(d1 > 0) || !(d2 > 0) || (!one_slice && !(dz > 0)))
        return fail(EMSIM_HIP_EINVAL, "slice_builder_create: voxel sizes must be positive");

    // Edge tables (float64 numpy.arange semantics, then rounded up to float32).
    const int ns[3] = {one_slice ? 1 : n_slices, n1, n2};
    const double ds[3] = {one_slice ? 1.0 : dz, d1, d2};
    std::vector<float> edges32;
    int last_exact[3];
    float inv_d[3];
    for (int a = 0; a < 3; ++a) {
        std::vector<double> e;
        reference_edges(ns[a], ds[a], e);
        if ((int)e.size() != ns[a] + 1) {
            if (a == 0 && one_slice) {
                e.resize(2);  // axis 0 is ignored by the one-slice builder
                e[0] = -1.0;
                e[1] = 1.0;
            } else {
                return fail(EMSIM_HIP_EUNSUPPORTED,
                            "binning axis %d: numpy.arange makes %d edges for %d bins of %.17g "
                            "(the reference would return a %d-bin histogram on this axis and "
                            "misalign its slices); choose a pixel size that divides evenly",
                            a, (int)e.size(), ns[a], ds[a], (int)e.size() - 1);
            }
        }
        for (double x : e) edges32.push_back(round_up_f32(x));
        last_exact[a] = ((double)(float)e.back() == e.back()) ? 1 : 0;
        inv_d[a] = (float)(1.0 / ds[a]);
    }

    emsim_hip_slice_builder* b = new (std::nothrow) emsim_hip_slice_builder();
    if (!b) return fail(EMSIM_HIP_ENOMEM, "slice_builder_create: out of host memory");
    b->n_elems = n_elems;
    b->n_slices = n_slices;
    b->n1 = n1;
    b->n2 = n2;
    b->H = n2 / 2 + 1;
    b->dz = dz;
    b->d1 = d1;
    b->d2 = d2;
    b->one_slice = one_slice ? 1 : 0;
    for (int a = 0; a < 3; ++a) {
        b->last_exact[a] = last_exact[a];
        b->inv_d[a] = inv_d[a];
    }
    int rows = kBinTileWords / n2;
    if (rows < 1) rows = 1;
    if (rows > n1) rows = n1;
    b->tile_rows = rows;
    b->tiles_per_plane = (int)ceil_div(n1, rows);

    int rc = get_twiddles(&b->tw);
    if (!rc && !(pow2_supported(n1) && pow2_supported(n2))) {
        b->chirp = true;
        rc = chirp_plan_create(n1, b->p1);
        if (!rc) rc = chirp_plan_create(n2, b->p2);
    }
    if (rc) {
        emsim_hip_slice_builder_destroy(b);
        return rc;
    }
    const size_t sf_bytes = (size_t)n_elems * n1 * b->H * sizeof(float);
    hipError_t err = hipMalloc(&b->sf, sf_bytes);
    if (err == hipSuccess) err = hipMemcpy(b->sf, sf_host, sf_bytes, hipMemcpyHostToDevice);
    if (err == hipSuccess && pow2_supported(n1)) {
        const int C = spec_tile_cols(n1), tiles = (int)ceil_div(b->H, C);
        std::vector<float> t((size_t)n_elems * tiles * n1 * C, 0.f);
        for (int e = 0; e < n_elems; ++e)
            for (int x = 0; x < n1; ++x)
                for (int j = 0; j < b->H; ++j)
                    t[(((size_t)e * tiles + j / C) * n1 + x) * C + j % C] =
                        sf_host[((size_t)e * n1 + x) * b->H + j];
        err = hipMalloc(&b->sf_tiles, t.size() * sizeof(float));
        if (err == hipSuccess)
            err = hipMemcpy(b->sf_tiles, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice);
    }
    if (err == hipSuccess) err = hipMalloc(&b->edges, edges32.size() * sizeof(float));
    if (err == hipSuccess)
        err = hipMemcpy(b->edges, edges32.data(), edges32.size() * sizeof(float),
                        hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        emsim_hip_slice_builder_destroy(b);
        return fail(EMSIM_HIP_EDEVICE, "slice_builder_create: %s", hipGetErrorString(err));
    }
    *out = b;
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_slice_builder_destroy(emsim_hip_slice_builder* b) {
    if (!b) return EMSIM_HIP_OK;
    chirp_plan_free(b->p1);
    chirp_plan_free(b->p2);
    if (b->sf) (void)hipFree(b->sf);
    if (b->sf_tiles) (void)hipFree(b->sf_tiles);
    if (b->edges) (void)hipFree(b->edges);
    delete b;
    return EMSIM_HIP_OK;
}

namespace {
struct BinWorkspace {
    int32_t* atom_tile;
    uint32_t* atom_local;
    uint32_t* bucket;
    uint32_t* tile_count;
    uint32_t* tile_start;
    uint32_t* block_sums;
    size_t bytes;
};

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

BinWorkspace bin_layout(const emsim_hip_slice_builder* b, int64_t n_atoms, char* base) {
    const int64_t n0 = b->one_slice ? 1 : b->n_slices;
    const int64_t n_tiles = (int64_t)b->n_elems * n0 * b->tiles_per_plane;
    const int64_t n_scan = n_tiles + 1;
    const int64_t n_blocks = ceil_div(n_scan, kScanTile);
    BinWorkspace w;
    size_t off = 0;
    w.atom_tile = (int32_t*)(base + off);  off += align_up(sizeof(int32_t) * (size_t)n_atoms);
    w.atom_local = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_atoms);
    w.bucket = (uint32_t*)(base + off);     off += align_up(sizeof(uint32_t) * (size_t)n_atoms);
    w.tile_count = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_scan);
    w.tile_start = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_scan);
    w.block_sums = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_blocks);
    w.bytes = off;
    return w;
}
}  // namespace

extern "C" size_t emsim_hip_bin_workspace_size(const emsim_hip_slice_builder* b, int64_t n_atoms) {
    if (!b || n_atoms < 0) return 0;
    return bin_layout(b, n_atoms, nullptr).bytes;
}

extern "C" int emsim_hip_bin_atoms(const emsim_hip_slice_builder* b, const float* coords,
                                   const int64_t* counts_host, int64_t n_atoms, float* hist,
                                   void* workspace, size_t workspace_bytes,
                                   emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!b || !hist || !counts_host || n_atoms < 0 || (n_atoms > 0 && !coords))
        return fail(EMSIM_HIP_EINVAL, "bin_atoms: bad arguments");
    if (n_atoms >= (int64_t)0x7fffffff)
        return fail(EMSIM_HIP_EUNSUPPORTED, "bin_atoms: at most 2^31-1 atoms per call");
    ElemOffsets eo;
    eo.n_elems = b->n_elems;
    eo.off[0] = 0;
    for (int e = 0; e < b->n_elems; ++e) {
        if (counts_host[e] < 0) return fail(EMSIM_HIP_EINVAL, "bin_atoms: negative count");
        eo.off[e + 1] = eo.off[e] + counts_host[e];
    }
    if (eo.off[b->n_elems] > n_atoms)
        return fail(EMSIM_HIP_EINVAL, "bin_atoms: counts sum to %lld but only %lld atoms given",
                    (long long)eo.off[b->n_elems], (long long)n_atoms);
    BinWorkspace w = bin_layout(b, n_atoms, (char*)workspace);
    if (!workspace || workspace_bytes < w.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "bin_atoms: workspace needs %zu bytes", w.bytes);

    const int64_t n0 = b->one_slice ? 1 : b->n_slices;
    BinGeom g;
    g.e0 = b->edges;
    g.e1 = b->edges + (n0 + 1);
    g.e2 = g.e1 + (b->n1 + 1);
    g.n0 = (int)n0;
    g.n1 = b->n1;
    g.n2 = b->n2;
    g.last0 = b->last_exact[0];
    g.last1 = b->last_exact[1];
    g.last2 = b->last_exact[2];
    g.inv0 = b->inv_d[0];
    g.inv1 = b->inv_d[1];
    g.inv2 = b->inv_d[2];
    g.one_slice = b->one_slice;
    g.tile_rows = b->tile_rows;
    g.tiles_per_plane = b->tiles_per_plane;

    const int64_t n_tiles = (int64_t)b->n_elems * n0 * b->tiles_per_plane;
    const int64_t n_scan = n_tiles + 1;
    const int64_t n_blocks = ceil_div(n_scan, kScanTile);

    EMSIM_LAUNCH(0, k_fill_u32, dim3(grid_for(n_scan, 256)), dim3(256), 0, s, w.tile_count,
                       n_scan, 0u);
    if (n_atoms > 0) {
        EMSIM_LAUNCH(0, k_bin_classify, dim3(grid_for(n_atoms, 256)), dim3(256), 0, s, coords,
                           n_atoms, eo, g, w.atom_tile, w.atom_local, w.tile_count);
    }
    EMSIM_LAUNCH(0, k_scan_tiles, dim3((unsigned)n_blocks), dim3(kScanBlock), 0, s,
                       w.tile_count, w.tile_start, n_scan, w.block_sums);
    EMSIM_LAUNCH(0, k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, w.block_sums, n_blocks);
    EMSIM_LAUNCH(0, k_scan_add, dim3((unsigned)n_blocks), dim3(kScanBlock), 0, s, w.tile_start,
                       n_scan, w.block_sums);
    if (n_atoms > 0) {
        EMSIM_LAUNCH(0, k_bin_scatter, dim3(grid_for(n_atoms, 256)), dim3(256), 0, s, n_atoms,
                           w.atom_tile, w.atom_local, w.tile_start, w.tile_count, w.bucket);
    }
    EMSIM_LAUNCH(0, k_bin_tiles, dim3((unsigned)n_tiles), dim3(kBinBlock), 0, s, w.bucket,
                       w.tile_start, g, hist);
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_add_water(const emsim_hip_slice_builder* b, float* hist, int idx_h,
                                   int idx_o, double voxel_water, uint64_t seed,
                                   emsim_stream_t stream_) {
    if (!b || !hist || idx_h < 0 || idx_o < 0 || idx_h >= b->n_elems || idx_o >= b->n_elems ||
        idx_h == idx_o || !(voxel_water >= 0))
        return fail(EMSIM_HIP_EINVAL, "add_water: bad arguments");
    const int64_t n0 = b->one_slice ? 1 : b->n_slices;
    const int64_t vox = n0 * b->n1 * b->n2;
    const float lam_o = (float)voxel_water, lam_h = (float)(2.0 * voxel_water);
    const float exp_o = (float)std::exp(-voxel_water), exp_h = (float)std::exp(-2.0 * voxel_water);
    EMSIM_LAUNCH(0, k_add_water, dim3(grid_for(vox, 256)), dim3(256), 0, (hipStream_t)stream_,
                       hist, b->n_elems, vox, idx_h, idx_o, lam_o, exp_o, lam_h, exp_h, seed);
    return EMSIM_HIP_OK;
}

namespace {
// Workspace of make_slices per slice of a chunk: the half spectra of every
// element, plus (chirp-z sizes) the line buffer of one element's transforms.
size_t slices_bytes_per_slice(const emsim_hip_slice_builder* b) {
    size_t per = (size_t)b->n_elems * b->n1 * half_ld(b->H) * sizeof(float2);
    if (b->chirp)
        per += sizeof(float2) * (size_t)std::max((int64_t)b->n1 * b->p2.M,
                                                 (int64_t)b->H * b->p1.M);
    return per;
}
}  // namespace

extern "C" size_t emsim_hip_slices_workspace_size(const emsim_hip_slice_builder* b,
                                                  int max_chunk) {
    if (!b) return 0;
    int chunk = (max_chunk <= 0 || max_chunk > b->n_slices) ? b->n_slices : max_chunk;
    return slices_bytes_per_slice(b) * (size_t)chunk;
}

extern "C" int emsim_hip_make_slices(const emsim_hip_slice_builder* b, const float* hist,
                                     float* out, int symmetric_bandlimit, void* workspace,
                                     size_t workspace_bytes, emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!b || !hist || !out) return fail(EMSIM_HIP_EINVAL, "make_slices: bad arguments");
    if (symmetric_bandlimit && b->n1 != b->n2)
        return fail(EMSIM_HIP_EINVAL,
                    "make_slices: the symmetric band-limit needs a square grid (the reference's "
                    "mask does not broadcast for %dx%d)", b->n1, b->n2);
    const int ldh = half_ld(b->H);
    const size_t per_slice = slices_bytes_per_slice(b);
    if (!workspace || workspace_bytes < per_slice)
        return fail(EMSIM_HIP_EWORKSPACE, "make_slices: workspace needs at least %zu bytes",
                    per_slice);
    int chunk = (int)(workspace_bytes / per_slice);
    if (chunk > b->n_slices) chunk = b->n_slices;
    if (chunk > 65535) chunk = 65535;
    float2* X = (float2*)workspace;
    const int r = (b->n1 < b->n2 ? b->n1 : b->n2) / 2;
    const float inv_npix = 1.0f / (float)((int64_t)b->n1 * b->n2);
    if (b->chirp) {
        // the same passes through chirp-z transforms: R2C rows, column DFTs,
        // spectrum sum, inverse column DFTs, C2R rows with the clip
        float2* work = X + (size_t)b->n_elems * chunk * b->n1 * ldh;
        const int n1 = b->n1, n2 = b->n2, H = b->H, nE = b->n_elems;
        for (int s0 = 0; s0 < b->n_slices; s0 += chunk) {
            const int sc = (b->n_slices - s0) < chunk ? (b->n_slices - s0) : chunk;
            for (int e = 0; e < nE; ++e) {
                float2* Xe = X + (size_t)e * sc * n1 * ldh;
                const float* in = hist + ((int64_t)e * b->n_slices + s0) * n1 * n2;
                int rc = line_dft(b->p2, false, kReal, in, rows_of(n2), (int64_t)sc * n1, Xe,
                                  rows_of(ldh), H, false, 0.f, work, b->tw, s);
                if (rc) return rc;
                // columns of sc images (n1 rows of ldh) of this element
                const LineMap cm{H, (int64_t)n1 * ldh, 1, ldh};
                rc = line_dft(b->p1, false, kComplex, Xe, cm, (int64_t)sc * H, Xe, cm, n1,
                              false, 0.f, work, b->tw, s);
                if (rc) return rc;
            }
            const int64_t total = (int64_t)sc * n1 * H;
            EMSIM_LAUNCH(0, k_half_spectrum_sum, dim3(grid_for(total, 256)), dim3(256), 0, s, X,
                         b->sf, nE, sc, n1, H, ldh, n2, symmetric_bandlimit ? 1 : 0, r * r,
                         inv_npix);
            const LineMap cm{H, (int64_t)n1 * ldh, 1, ldh};
            int rc = line_dft(b->p1, true, kComplex, X, cm, (int64_t)sc * H, X, cm, n1, false,
                              0.f, work, b->tw, s);
            if (rc) return rc;
            rc = line_dft(b->p2, true, kHalf, X, rows_of(ldh), (int64_t)sc * n1,
                          out + (int64_t)s0 * n1 * n2, rows_of(n2), n2, true, 1e-13f, work,
                          b->tw, s);
            if (rc) return rc;
        }
        return EMSIM_HIP_OK;
    }
    for (int s0 = 0; s0 < b->n_slices; s0 += chunk) {
        const int sc = (b->n_slices - s0) < chunk ? (b->n_slices - s0) : chunk;
        int rc = r2c_pairs(b->n2, hist, X, (int64_t)b->n_elems * sc * (b->n1 / 2), b->n1, sc,
                           b->n_slices, s0, ldh, b->tw, s);
        if (rc) return rc;
        rc = col_spectrum(b->n1, X, b->sf, b->n_elems, sc, b->H, ldh, b->n2,
                          symmetric_bandlimit ? 1 : 0, r * r, inv_npix, b->tw, s);
        if (rc) return rc;
        rc = c2r_pairs(b->n2, X, out + (int64_t)s0 * b->n1 * b->n2, (int64_t)sc * (b->n1 / 2),
                       1e-13f, ldh, b->tw, s);
        if (rc) return rc;
    }
    return EMSIM_HIP_OK;
}

// ==========================================================================
// wave propagator
// ==========================================================================
extern "C" int emsim_hip_wave_propagator_create(emsim_hip_wave_propagator** out, int n1, int n2,
                                                float d1, float d2, float wave_length,
                                                float relativity_gamma) {
    if (!out) return fail(EMSIM_HIP_EINVAL, "wave_propagator_create: null output handle");
    *out = nullptr;
    if (!size_supported(n1) || !size_supported(n2))
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "wave_propagator_create: wave shape %dx%d; the hip backend implements "
                    "sizes from 8 to %d (powers of two up to %d)", n1, n2, kChirpMaxN, kTwN);
    if (!(d1 > 0) || !(d2 > 0) || !(wave_length > 0))
        return fail(EMSIM_HIP_EINVAL, "wave_propagator_create: sizes must be positive");
    emsim_hip_wave_propagator* w = new (std::nothrow) emsim_hip_wave_propagator();
    if (!w) return fail(EMSIM_HIP_ENOMEM, "wave_propagator_create: out of host memory");
    w->n1 = n1;
    w->n2 = n2;
    w->d1 = d1;
    w->d2 = d2;
    w->lam = wave_length;
    w->gamma = relativity_gamma;
    // WavePropagator.cpp:19-23
    float dfx = 1.0f / d1;
    w->dfx = dfx / (float)n1;
    float dfy = 1.0f / d2;
    w->dfy = dfy / (float)n2;
    const float fmax = 0.5f / (d1 >= d2 ? d1 : d2);
    const float filt = 0.6667f * fmax;
    w->filter2 = filt * filt;
    w->factor = wave_length * relativity_gamma;   // WavePropagator.cpp:42
    int rc = get_twiddles(&w->tw);
    if (!rc && !(pow2_supported(n1) && pow2_supported(n2))) {
        w->chirp = true;
        rc = chirp_plan_create(n1, w->p1);
        if (!rc) rc = chirp_plan_create(n2, w->p2);
        if (!rc) {
            const int64_t lines = std::max((int64_t)n1 * w->p2.M, (int64_t)n2 * w->p1.M);
            if (hipMalloc(&w->work, sizeof(float2) * (size_t)lines) != hipSuccess)
                rc = fail(EMSIM_HIP_ENOMEM, "wave_propagator_create: no device memory for "
                                            "the chirp-z line buffer");
        }
    }
    if (rc) {
        emsim_hip_wave_propagator_destroy(w);
        return rc;
    }
    *out = w;
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_wave_propagator_destroy(emsim_hip_wave_propagator* w) {
    if (!w) return EMSIM_HIP_OK;
    chirp_plan_free(w->p1);
    chirp_plan_free(w->p2);
    if (w->work) (void)hipFree(w->work);
    delete w;
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_init_wave(const emsim_hip_wave_propagator* w, float value, float* wave,
                                   int batch, emsim_stream_t stream_) {
    if (!w || !wave || batch < 1) return fail(EMSIM_HIP_EINVAL, "init_wave: bad arguments");
    const int64_t n = (int64_t)batch * w->n1 * w->n2;
    EMSIM_LAUNCH(0, k_fill_complex, dim3(grid_for(n, 256)), dim3(256), 0,
                       (hipStream_t)stream_, (float2*)wave, n, value);
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_slice_transmit(const emsim_hip_wave_propagator* w, const float* wave,
                                        const float* slice, int64_t slice_batch_stride,
                                        float* out, int batch, emsim_stream_t stream_) {
    if (!w || !wave || !slice || !out || batch < 1 || slice_batch_stride < 0)
        return fail(EMSIM_HIP_EINVAL, "slice_transmit: bad arguments");
    const int64_t npix = (int64_t)w->n1 * w->n2;
    const int64_t n = npix * batch;
    EMSIM_LAUNCH(0, k_transmit, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream_,
                       (const float2*)wave, slice, (float2*)out, npix, n, slice_batch_stride,
                       w->factor);
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_space_propagate(const emsim_hip_wave_propagator* w, const float* wave,
                                         float dz, float* out, int batch,
                                         emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!w || !wave || !out || batch < 1) return fail(EMSIM_HIP_EINVAL, "space_propagate: bad arguments");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "space_propagate: batch <= 65535");
    if (w->chirp)
        return chirp_filter(w, (const float2*)wave, (float2*)out, batch, make_propagator(w, dz), s);
    const int64_t rows = (int64_t)batch * w->n1;
    const int64_t npix = (int64_t)w->n1 * w->n2;
    float2* o = (float2*)out;
    int rc = row_fft(w->n2, false, (const float2*)wave, o, rows, w->tw, s);
    if (!rc) rc = col_prop(w->n1, o, w->n2, w->n2, npix, batch, make_propagator(w, dz), w->tw, s);
    if (!rc) rc = row_fft(w->n2, true, o, o, rows, w->tw, s);
    return rc;
}

extern "C" int emsim_hip_multislice_propagate(const emsim_hip_wave_propagator* w,
                                              const float* wave, const float* slices,
                                              int n_slices, int64_t slices_batch_stride,
                                              float dz, float* out, int batch,
                                              emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!w || !wave || !slices || !out || batch < 1 || n_slices < 1 || slices_batch_stride < 0)
        return fail(EMSIM_HIP_EINVAL, "multislice_propagate: bad arguments");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "multislice_propagate: batch <= 65535");
    const int64_t rows = (int64_t)batch * w->n1;
    const int64_t npix = (int64_t)w->n1 * w->n2;
    const PropagatorMul prop = make_propagator(w, dz);
    float2* o = (float2*)out;
    if (w->chirp) {
        // image by image: transmission, then the propagator through chirp-z
        // transforms (WavePropagator.cpp:90-103)
        for (int b = 0; b < batch; ++b) {
            float2* ob = o + b * npix;
            for (int sl = 0; sl < n_slices; ++sl) {
                const float2* src = sl == 0 ? (const float2*)wave + b * npix : ob;
                EMSIM_LAUNCH(0, k_transmit, dim3(grid_for(npix, 256)), dim3(256), 0, s, src,
                             slices + b * slices_batch_stride + (int64_t)sl * npix, ob, npix, npix,
                             (int64_t)0, w->factor);
                const int rc = chirp_filter(w, ob, ob, 1, prop, s);
                if (rc) return rc;
            }
        }
        return EMSIM_HIP_OK;
    }
    for (int sl = 0; sl < n_slices; ++sl) {
        const float2* src = sl == 0 ? (const float2*)wave : o;
        int rc = row_transmit_fft(w->n2, sl == 0, src, slices + (int64_t)sl * npix, o, rows,
                                  w->n1, slices_batch_stride, w->factor, w->tw, s);
        if (rc) return rc;
        rc = col_prop(w->n1, o, w->n2, w->n2, npix, batch, prop, w->tw, s);
        if (rc) return rc;
    }
    return row_fft(w->n2, true, o, o, rows, w->tw, s);
}

extern "C" int emsim_hip_lens_propagate(const emsim_hip_wave_propagator* w, const float* wave,
                                        float cs_mm, float defocus, float aperture, float* out,
                                        int batch, emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!w || !wave || !out || batch < 1) return fail(EMSIM_HIP_EINVAL, "lens_propagate: bad arguments");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "lens_propagate: batch <= 65535");
    if (w->chirp)
        return chirp_filter(w, (const float2*)wave, (float2*)out, batch,
                            make_lens(w, cs_mm, defocus, aperture), s);
    const int64_t rows = (int64_t)batch * w->n1;
    const int64_t npix = (int64_t)w->n1 * w->n2;
    float2* o = (float2*)out;
    int rc = row_fft(w->n2, false, (const float2*)wave, o, rows, w->tw, s);
    if (!rc)
        rc = col_lens(w->n1, o, w->n2, w->n2, npix, batch, make_lens(w, cs_mm, defocus, aperture),
                      w->tw, s);
    if (!rc) rc = row_fft(w->n2, true, o, o, rows, w->tw, s);
    return rc;
}

extern "C" int emsim_hip_lens_intensity(const emsim_hip_wave_propagator* w, const float* wave,
                                        float cs_mm, float defocus, float aperture,
                                        float* scratch, float* image, int batch,
                                        emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!w || !wave || !scratch || !image || batch < 1)
        return fail(EMSIM_HIP_EINVAL, "lens_intensity: bad arguments");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "lens_intensity: batch <= 65535");
    const int64_t rows = (int64_t)batch * w->n1;
    const int64_t npix = (int64_t)w->n1 * w->n2;
    float2* o = (float2*)scratch;
    if (w->chirp) {
        const int rc = chirp_filter(w, (const float2*)wave, o, batch,
                                    make_lens(w, cs_mm, defocus, aperture), s);
        if (rc) return rc;
        EMSIM_LAUNCH(0, k_intensity, dim3(grid_for(npix * batch, 256)), dim3(256), 0, s,
                     (const float2*)o, image, npix * batch);
        return EMSIM_HIP_OK;
    }
    int rc = row_fft(w->n2, false, (const float2*)wave, o, rows, w->tw, s);
    if (!rc)
        rc = col_lens(w->n1, o, w->n2, w->n2, npix, batch, make_lens(w, cs_mm, defocus, aperture),
                      w->tw, s);
    if (!rc) rc = row_ifft_intensity(w->n2, o, image, rows, w->tw, s);
    return rc;
}

extern "C" int emsim_hip_intensity(const float* wave, float* out, int64_t n,
                                   emsim_stream_t stream_) {
    if (!wave || !out || n < 0) return fail(EMSIM_HIP_EINVAL, "intensity: bad arguments");
    if (n == 0) return EMSIM_HIP_OK;
    EMSIM_LAUNCH(0, k_intensity, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream_,
                       (const float2*)wave, out, n);
    return EMSIM_HIP_OK;
}

// ==========================================================================
// per-kernel timing
// ==========================================================================
extern "C" int emsim_hip_timing_enable(int on) {
    std::lock_guard<std::mutex> lock(g_timing_mutex);
    g_timing_on = on != 0;
    return EMSIM_HIP_OK;
}

// ==========================================================================
// one molecule in many orientations, on the device
// ==========================================================================
extern "C" size_t emsim_hip_orient_workspace_size(int batch) {
    return batch < 1 ? 0 : align_up(sizeof(uint32_t) * 6 * (size_t)batch);
}

namespace {
// The three launches of an orientation batch; bbox holds 6 keys per instance
// (device memory, cleared here by a fill kernel before use).
int orient_launch(const float* base, int64_t n_atoms, const float* rot, float c0x, float c0y,
                  float c0z, int batch, float* out, uint32_t* bbox, hipStream_t s) {
    EMSIM_LAUNCH(0, k_bbox_init, dim3((unsigned)ceil_div(6 * (int64_t)batch, 256)), dim3(256), 0,
                 s, bbox, batch);
    EMSIM_LAUNCH(0, k_orient, dim3((unsigned)ceil_div(n_atoms, 256), (unsigned)batch), dim3(256),
                 0, s, base, n_atoms, rot, c0x, c0y, c0z, out, bbox);
    const int64_t total = (int64_t)batch * n_atoms * 3;
    EMSIM_LAUNCH(0, k_recenter, dim3(grid_for(total, 256)), dim3(256), 0, s, out, n_atoms, batch,
                 (const uint32_t*)bbox);
    return EMSIM_HIP_OK;
}
}  // namespace

extern "C" int emsim_hip_orient_batch(const float* base, int64_t n_atoms, const float* rot,
                                      float c0x, float c0y, float c0z, int batch, float* out,
                                      void* workspace, size_t workspace_bytes,
                                      emsim_stream_t stream_) {
    if (!base || !rot || !out || !workspace || n_atoms < 1 || batch < 1)
        return fail(EMSIM_HIP_EINVAL, "orient_batch: bad arguments");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "orient_batch: batch <= 65535");
    if (workspace_bytes < emsim_hip_orient_workspace_size(batch))
        return fail(EMSIM_HIP_EWORKSPACE, "orient_batch: workspace too small");
    return orient_launch(base, n_atoms, rot, c0x, c0y, c0z, batch, out, (uint32_t*)workspace,
                         (hipStream_t)stream_);
}

// --------------------------------------------------------------------------
// the same without a workspace argument, for callers that cannot allocate
// --------------------------------------------------------------------------
// The bounding-box keys live in a library-owned device array.  In host code
// `scratch` only names the symbol; its device address (one copy per GPU) comes
// from hipGetSymbolAddress, taken once per device with that device current.
static __device__ __align__(256) uint32_t scratch[6 * EMSIM_HIP_ORIENT_NOALLOC_MAX_BATCH];

namespace {
std::mutex g_scratch_mutex;
std::vector<uint32_t*> g_scratch_ptr;   // per device; null until first use

int scratch_address(uint32_t** out) {
    int dev = 0;
    EMSIM_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (err != hipSuccess || !p)
            return fail(EMSIM_HIP_EDEVICE, "orient_batch_noalloc: hipGetSymbolAddress failed: %s",
                        hipGetErrorString(err));
        g_scratch_ptr[dev] = (uint32_t*)p;
    }
    *out = g_scratch_ptr[dev];
    return EMSIM_HIP_OK;
}
}  // namespace

extern "C" int emsim_hip_orient_batch_noalloc(const float* base, int64_t n_atoms,
                                              const float* rot, float c0x, float c0y, float c0z,
                                              int batch, float* out, emsim_stream_t stream_) {
    if (!base || !rot || !out || n_atoms < 1 || batch < 1)
        return fail(EMSIM_HIP_EINVAL, "orient_batch_noalloc: bad arguments");
    if (batch > EMSIM_HIP_ORIENT_NOALLOC_MAX_BATCH)
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "orient_batch_noalloc: batch %d exceeds the library's scratch (at most %d "
                    "orientations per call; use emsim_hip_orient_batch with a workspace)",
                    batch, EMSIM_HIP_ORIENT_NOALLOC_MAX_BATCH);
    uint32_t* bbox = nullptr;
    const int rc = scratch_address(&bbox);
    if (rc) return rc;
    return orient_launch(base, n_atoms, rot, c0x, c0y, c0z, batch, out, bbox,
                         (hipStream_t)stream_);
}

// ==========================================================================
// launch-shape knobs
// ==========================================================================
extern "C" int emsim_hip_tuning_set(const char* key, int value) {
    if (!key) return fail(EMSIM_HIP_EINVAL, "tuning_set: null key");
    if (std::strcmp(key, "row_items_per_block") == 0) {
        if (value < 0 || value > 64) return fail(EMSIM_HIP_EINVAL, "row_items_per_block in [0, 64]");
        g_row_items_per_block = value;
    } else if (std::strcmp(key, "col_tiles_per_block") == 0) {
        if (value < 0 || value > 32) return fail(EMSIM_HIP_EINVAL, "col_tiles_per_block in [0, 32]");
        g_col_tiles_per_block = value;
    } else if (std::strcmp(key, "col_shape") == 0) {
        if (value < 0 || value > 2) return fail(EMSIM_HIP_EINVAL, "col_shape in [0, 2]");
        g_col_shape = value;
    } else {
        return fail(EMSIM_HIP_EINVAL, "tuning_set: unknown key '%s'", key);
    }
    return EMSIM_HIP_OK;
}

extern "C" int emsim_hip_tuning_get(const char* key) {
    if (key && std::strcmp(key, "row_items_per_block") == 0) return g_row_items_per_block;
    if (key && std::strcmp(key, "col_tiles_per_block") == 0) return g_col_tiles_per_block;
    if (key && std::strcmp(key, "col_shape") == 0) return g_col_shape;
    return fail(EMSIM_HIP_EINVAL, "tuning_get: unknown key");
}

extern "C" int emsim_hip_timing_collect(char* buf, size_t cap, size_t* needed) {
    struct Agg {
        std::string name;
        int n;
        long count;
        double total, lo, hi;
    };
    std::vector<Agg> aggs;
    std::vector<TimingRecord> recs;
    {
        std::lock_guard<std::mutex> lock(g_timing_mutex);
        recs.swap(g_timing_records);
    }
    int rc = EMSIM_HIP_OK;
    for (const TimingRecord& r : recs) {
        float ms = 0.f;
        hipError_t err = hipEventSynchronize(r.stop);
        if (err == hipSuccess) err = hipEventElapsedTime(&ms, r.start, r.stop);
        if (err != hipSuccess) {
            rc = fail(EMSIM_HIP_EDEVICE, "timing_collect: %s", hipGetErrorString(err));
            continue;
        }
        Agg* a = nullptr;
        for (Agg& x : aggs)
            if (x.n == r.n && x.name == r.name) a = &x;
        if (!a) {
            aggs.push_back({r.name, r.n, 0, 0.0, 1e30, 0.0});
            a = &aggs.back();
        }
        a->count += 1;
        a->total += ms;
        a->lo = ms < a->lo ? ms : a->lo;
        a->hi = ms > a->hi ? ms : a->hi;
    }
    {
        std::lock_guard<std::mutex> lock(g_timing_mutex);
        for (const TimingRecord& r : recs) {
            g_event_pool.push_back(r.start);
            g_event_pool.push_back(r.stop);
        }
    }
    std::string text;
    char line[256];
    for (const Agg& a : aggs) {
        snprintf(line, sizeof(line), "%s\t%d\t%ld\t%.6f\t%.6f\t%.6f\n", a.name.c_str(), a.n,
                 a.count, a.total, a.lo, a.hi);
        text += line;
    }
    if (needed) *needed = text.size() + 1;
    if (buf && cap > 0) {
        const size_t m = text.size() < cap - 1 ? text.size() : cap - 1;
        memcpy(buf, text.data(), m);
        buf[m] = '\0';
    }
    return rc;
}

// ==========================================================================
// fused batched path (kernels_fused.hpp)
// ==========================================================================
namespace {

template <int N>
int launch_spec_sparse(const uint32_t* row_start, const uint16_t* bucket, const float* sf,
                       const float* sft, float2* Y, int slice, int n_slices, int n_elems, int H,
                       int ldh, int n2, int bandlimit, int band_r2, float inv_npix, int batch,
                       const int* order, const float2* tw, hipStream_t s) {
    using Cfg = SpecCfg<N>;
    if (H != n2 / 2 + 1 || n2 < 2 || ldh < H + n_elems)
        return fail(EMSIM_HIP_EINVAL, "spectrum: needs H = n2/2 + 1, n2 >= 2, ldh >= H + n_elems");
    // with spec_pack(N), column n2/2 is carried by column 0
    const int cols = spec_pack(N) ? n2 / 2 : H;
    const dim3 grid((unsigned)(ceil_div(cols, Cfg::C) * (int64_t)batch));
    if (!sft) return fail(EMSIM_HIP_EINVAL, "spectrum: no tile-major form-factor table");
    EMSIM_LAUNCH(N, (k_spec_sparse<N>), grid, dim3(Cfg::BLOCK), 0, s, row_start, bucket, sf,
                 sft, Y, slice, n_slices, n_elems, H, ldh, n2, batch, bandlimit, band_r2,
                 inv_npix, order, tw);
    return EMSIM_HIP_OK;
}

// vout (nullable) receives the slice's potential, (B, n_slices, n1, N)
template <int N>
int launch_slice_pass_a(bool first, const float2* Y, float2* wave, int64_t n_pairs, int ldh,
                        float psi0, float factor, float* vout, int slice, int n_slices, int n1,
                        ColDead cd, const float2* tw, hipStream_t s) {
    // one round of resident blocks: pass A is latency-bound, and a partial
    // second round (or a grid below the resident count) leaves SIMDs idle
    // The row-pair groups are dealt out evenly over one round of resident
    // waves (RunSched), so every SIMD ends up with about the same number of
    // pairs; the rows-per-block knob instead fixes the blocks' count.
    static int resident_cache[64];
    const int64_t resident =
        resident_blocks(k_slice_pass_a<N, false>, RowCfg<N>::BLOCK, resident_cache);
    const int64_t groups = ceil_div(n_pairs, RowCfg<N>::G);
    if (groups <= 0) return EMSIM_HIP_OK;
    int64_t blocks;
    if (g_row_items_per_block > 0) blocks = ceil_div(groups, g_row_items_per_block);
    else if (resident > 0) blocks = groups < resident ? groups : resident;
    else blocks = ceil_div(groups, rows_per_block(groups));
    const RunSched sched{(int)(groups / blocks), (int)(groups % blocks)};
    const dim3 grid((unsigned)blocks);
    const dim3 block(RowCfg<N>::BLOCK);
    const int64_t plane = (int64_t)n1 * N;
    if (first)
        EMSIM_LAUNCH(N, (k_slice_pass_a<N, true>), grid, block, 0, s,
                     (OpSlicePassA<N, true>{Y, wave, ldh, psi0, factor, vout, n_slices * plane,
                                            slice * plane, n1, cd.lo, cd.hi}),
                     n_pairs, sched, tw);
    else
        EMSIM_LAUNCH(N, (k_slice_pass_a<N, false>), grid, block, 0, s,
                     (OpSlicePassA<N, false>{Y, wave, ldh, psi0, factor, vout, n_slices * plane,
                                             slice * plane, n1, cd.lo, cd.hi}),
                     n_pairs, sched, tw);
    return EMSIM_HIP_OK;
}

template <int N>
int launch_col_lensprop(float2* data, int ld, int ncols, int64_t img_stride, int n_img,
                        LensPropMul mul, const float2* tw, hipStream_t s, ColDead cd) {
    return launch_col_mul_impl<N>(data, ld, ncols, img_stride, n_img, mul, tw, s, cd, true);
}

int spec_sparse(int n, const uint32_t* row_start, const uint16_t* bucket, const float* sf,
                const float* sft, float2* Y, int slice, int n_slices, int n_elems, int H, int ldh,
                int n2, int bandlimit, int band_r2, float inv_npix, int batch, const int* order,
                const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_spec_sparse, row_start, bucket, sf, sft, Y, slice, n_slices,
                   n_elems, H, ldh, n2, bandlimit, band_r2, inv_npix, batch, order, tw, s);
}
int slice_pass_a(int n, bool first, const float2* Y, float2* wave, int64_t n_pairs, int ldh,
                 float psi0, float factor, float* vout, int slice, int n_slices, int n1,
                 ColDead cd, const float2* tw, hipStream_t s) {
    EMSIM_DISPATCH(n, launch_slice_pass_a, first, Y, wave, n_pairs, ldh, psi0, factor, vout,
                   slice, n_slices, n1, cd, tw, s);
}
int col_lensprop(int n, float2* data, int ld, int ncols, int64_t img_stride, int n_img,
                 LensPropMul mul, const float2* tw, hipStream_t s, ColDead cd) {
    EMSIM_DISPATCH(n, launch_col_lensprop, data, ld, ncols, img_stride, n_img, mul, tw, s, cd);
}

struct FusedWorkspace {
    int32_t* order;   // spectrum-pass molecule slots per slice, or null (identity)
    int64_t* mol_off;
    int64_t* elem_off;
    int32_t* atom_row;
    uint16_t* atom_y;
    uint16_t* bucket;
    uint32_t* row_count;
    uint32_t* row_start;
    uint32_t* block_sums;
    float2* Y;
    float2* wave;
    // water fill (emsim_hip_simulate_water): dense count planes and half
    // spectra of one slice of every molecule, per-molecule seeds
    float* hist;
    float2* X;
    uint64_t* seeds;
    // molecules given as they come (emsim_hip_simulate_raw): per-atom element
    // indices and bounding-box keys (batch, 6)
    const uint8_t* atom_elem;
    uint32_t* bbox;
    int64_t max_mol_atoms;
    int64_t n_rows;
    size_t bytes;
};

// Water-fill parameters of a fused run (emsim_hip_simulate_water).
struct WaterFill {
    bool on = false;
    int idx_h = 0, idx_o = 0;
    float lam_o = 0.f, exp_o = 1.f, lam_h = 0.f, exp_h = 1.f;
};

FusedWorkspace fused_layout(const emsim_hip_slice_builder* b, int batch, int64_t n_atoms,
                            bool own_wave, char* base, bool water = false, bool raw = false) {
    FusedWorkspace w;
    w.n_rows = (int64_t)batch * b->n_slices * b->n_elems * b->n1;
    const int64_t n_scan = w.n_rows + 1;
    const int64_t n_blocks = ceil_div(n_scan, kScanTile);
    const int ldh = fused_ld(b);
    size_t off = 0;
    w.order = (int32_t*)(base + off);
    off += align_up(sizeof(int32_t) * (size_t)batch * b->n_slices);
    w.mol_off = (int64_t*)(base + off);    off += align_up(sizeof(int64_t) * (size_t)(batch + 1));
    w.elem_off = (int64_t*)(base + off);
    off += align_up(sizeof(int64_t) * (size_t)batch * (b->n_elems + 1));
    w.atom_row = (int32_t*)(base + off);   off += align_up(sizeof(int32_t) * (size_t)n_atoms);
    w.atom_y = (uint16_t*)(base + off);    off += align_up(sizeof(uint16_t) * (size_t)n_atoms);
    w.bucket = (uint16_t*)(base + off);    off += align_up(sizeof(uint16_t) * (size_t)n_atoms);
    w.row_count = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_scan);
    w.row_start = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_scan);
    w.block_sums = (uint32_t*)(base + off); off += align_up(sizeof(uint32_t) * (size_t)n_blocks);
    w.Y = (float2*)(base + off);
    off += align_up(sizeof(float2) * (size_t)batch * b->n1 * ldh);
    w.wave = nullptr;
    if (own_wave) {
        w.wave = (float2*)(base + off);
        off += align_up(sizeof(float2) * (size_t)batch * b->n1 * b->n2);
    }
    w.hist = nullptr;
    w.X = nullptr;
    w.seeds = nullptr;
    if (water) {
        const size_t planes = (size_t)b->n_elems * batch * b->n1;
        w.hist = (float*)(base + off);   off += align_up(sizeof(float) * planes * b->n2);
        w.X = (float2*)(base + off);     off += align_up(sizeof(float2) * planes * half_ld(b->H));
        w.seeds = (uint64_t*)(base + off); off += align_up(sizeof(uint64_t) * (size_t)batch);
    }
    w.atom_elem = nullptr;
    w.bbox = nullptr;
    w.max_mol_atoms = 0;
    if (raw) {
        w.bbox = (uint32_t*)(base + off);
        off += align_up(sizeof(uint32_t) * 6 * (size_t)batch);
    }
    w.bytes = off;
    return w;
}

int check_simulate(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                   const float* coords, int64_t n_atoms, int batch, int symmetric_bandlimit,
                   const float* out, int mode) {
    if (!b || !w || !out || batch < 1 || n_atoms < 0 || (n_atoms > 0 && !coords) ||
        (mode != 0 && mode != 1))
        return fail(EMSIM_HIP_EINVAL, "simulate: bad arguments");
    if (w->n1 != b->n1 || w->n2 != b->n2)
        return fail(EMSIM_HIP_EINVAL, "simulate: builder grid %dx%d != propagator grid %dx%d",
                    b->n1, b->n2, w->n1, w->n2);
    if (b->chirp || w->chirp)
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "simulate: the batched fused path needs power-of-two lateral sizes (%dx%d); "
                    "other sizes run through the per-call API", b->n1, b->n2);
    if (symmetric_bandlimit && b->n1 != b->n2)
        return fail(EMSIM_HIP_EINVAL, "simulate: the symmetric band-limit needs a square grid");
    if (batch > 65535) return fail(EMSIM_HIP_EUNSUPPORTED, "simulate: batch <= 65535");
    if (n_atoms >= (int64_t)0x7fffffff)
        return fail(EMSIM_HIP_EUNSUPPORTED, "simulate: at most 2^31-1 atoms per call");
    const int64_t n_rows_total = (int64_t)batch * b->n_slices * b->n_elems * b->n1;
    if (n_rows_total >= (int64_t)0x7fffffff)
        return fail(EMSIM_HIP_EUNSUPPORTED, "simulate: batch * slices * elements * n1 too large");
    return EMSIM_HIP_OK;
}

int simulate_core(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                  const float* coords, int64_t n_atoms, const FusedWorkspace& ws, int batch,
                  float psi0, float cs_mm, float defocus, float aperture, int symmetric_bandlimit,
                  float* out, float* slices_out, int mode, hipStream_t s,
                  const WaterFill& water = WaterFill{});

// Host-side offsets of emsim_hip_simulate: elem_off (batch, n_elems + 1).
int host_offsets(const emsim_hip_slice_builder* b, const int64_t* mol_off_host,
                 const int64_t* elem_counts_host, int batch, int64_t n_atoms,
                 std::vector<int64_t>& elem_off) {
    if (mol_off_host[0] != 0 || mol_off_host[batch] != n_atoms)
        return fail(EMSIM_HIP_EINVAL, "simulate: molecule offsets must run from 0 to n_atoms");
    elem_off.assign((size_t)batch * (b->n_elems + 1), 0);
    for (int m = 0; m < batch; ++m) {
        if (mol_off_host[m + 1] < mol_off_host[m])
            return fail(EMSIM_HIP_EINVAL, "simulate: molecule offsets must not decrease");
        int64_t o = mol_off_host[m];
        elem_off[(size_t)m * (b->n_elems + 1)] = o;
        for (int e = 0; e < b->n_elems; ++e) {
            const int64_t c = elem_counts_host[(size_t)m * b->n_elems + e];
            if (c < 0) return fail(EMSIM_HIP_EINVAL, "simulate: negative element count");
            o += c;
            elem_off[(size_t)m * (b->n_elems + 1) + e + 1] = o;
        }
        if (o > mol_off_host[m + 1])
            return fail(EMSIM_HIP_EINVAL, "simulate: molecule %d's element counts exceed its atoms",
                        m);
    }
    return EMSIM_HIP_OK;
}

}  // namespace

extern "C" size_t emsim_hip_simulate_workspace_size(const emsim_hip_slice_builder* b, int batch,
                                                    int64_t n_atoms, int mode) {
    if (!b || batch < 1 || n_atoms < 0) return 0;
    return fused_layout(b, batch, n_atoms, mode == 0, nullptr).bytes;
}

extern "C" int emsim_hip_simulate(const emsim_hip_slice_builder* b,
                                  const emsim_hip_wave_propagator* w, const float* coords,
                                  int64_t n_atoms, const int64_t* mol_off_host,
                                  const int64_t* elem_counts_host, int batch, float psi0,
                                  float cs_mm, float defocus, float aperture,
                                  int symmetric_bandlimit, float* out, float* slices_out,
                                  int mode, void* workspace, size_t workspace_bytes,
                                  emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!mol_off_host || !elem_counts_host) return fail(EMSIM_HIP_EINVAL, "simulate: null offsets");
    const int check = check_simulate(b, w, coords, n_atoms, batch, symmetric_bandlimit, out, mode);
    if (check) return check;
    std::vector<int64_t> elem_off;
    const int rc = host_offsets(b, mol_off_host, elem_counts_host, batch, n_atoms, elem_off);
    if (rc) return rc;
    FusedWorkspace ws = fused_layout(b, batch, n_atoms, mode == 0, (char*)workspace);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "simulate: workspace needs %zu bytes", ws.bytes);
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.mol_off, mol_off_host, sizeof(int64_t) * (batch + 1),
                                   hipMemcpyHostToDevice, s));
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.elem_off, elem_off.data(), sizeof(int64_t) * elem_off.size(),
                                   hipMemcpyHostToDevice, s));
    return simulate_core(b, w, coords, n_atoms, ws, batch, psi0, cs_mm, defocus, aperture,
                         symmetric_bandlimit, out, slices_out, mode, s);
}

extern "C" int emsim_hip_simulate_dev(const emsim_hip_slice_builder* b,
                                      const emsim_hip_wave_propagator* w, const float* coords,
                                      int64_t n_atoms, const int64_t* mol_off_dev,
                                      const int64_t* elem_off_dev, int batch, float psi0,
                                      float cs_mm, float defocus, float aperture,
                                      int symmetric_bandlimit, float* out, float* slices_out,
                                      int mode, void* workspace, size_t workspace_bytes,
                                      emsim_stream_t stream_) {
    if (!mol_off_dev || !elem_off_dev) return fail(EMSIM_HIP_EINVAL, "simulate_dev: null offsets");
    int rc = check_simulate(b, w, coords, n_atoms, batch, symmetric_bandlimit, out, mode);
    if (rc) return rc;
    FusedWorkspace ws = fused_layout(b, batch, n_atoms, mode == 0, (char*)workspace);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "simulate: workspace needs %zu bytes", ws.bytes);
    ws.mol_off = const_cast<int64_t*>(mol_off_dev);
    ws.elem_off = const_cast<int64_t*>(elem_off_dev);
    return simulate_core(b, w, coords, n_atoms, ws, batch, psi0, cs_mm, defocus, aperture,
                         symmetric_bandlimit, out, slices_out, mode, (hipStream_t)stream_);
}

extern "C" size_t emsim_hip_simulate_water_workspace_size(const emsim_hip_slice_builder* b,
                                                          int batch, int64_t n_atoms, int mode) {
    if (!b || batch < 1 || n_atoms < 0) return 0;
    return fused_layout(b, batch, n_atoms, mode == 0, nullptr, true).bytes;
}

extern "C" int emsim_hip_simulate_water(const emsim_hip_slice_builder* b,
                                        const emsim_hip_wave_propagator* w, const float* coords,
                                        int64_t n_atoms, const int64_t* mol_off_host,
                                        const int64_t* elem_counts_host, int batch, float psi0,
                                        float cs_mm, float defocus, float aperture,
                                        int symmetric_bandlimit, int idx_h, int idx_o,
                                        double voxel_water, const uint64_t* seeds_host,
                                        float* out, float* slices_out, int mode, void* workspace,
                                        size_t workspace_bytes, emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!mol_off_host || !elem_counts_host || !seeds_host)
        return fail(EMSIM_HIP_EINVAL, "simulate_water: null offsets or seeds");
    int rc = check_simulate(b, w, coords, n_atoms, batch, symmetric_bandlimit, out, mode);
    if (rc) return rc;
    if (b->one_slice)
        return fail(EMSIM_HIP_EINVAL, "simulate_water: needs a multi-slice builder (the "
                    "one-slice path adds no water)");
    if (idx_h < 0 || idx_o < 0 || idx_h >= b->n_elems || idx_o >= b->n_elems || idx_h == idx_o ||
        !(voxel_water >= 0))
        return fail(EMSIM_HIP_EINVAL, "simulate_water: bad water arguments");
    if ((size_t)b->n_elems * b->n2 * sizeof(float) > 64 * 1024)
        return fail(EMSIM_HIP_EUNSUPPORTED,
                    "simulate_water: n_elems * n2 <= 16384 (one row of every element in LDS)");
    std::vector<int64_t> elem_off;
    rc = host_offsets(b, mol_off_host, elem_counts_host, batch, n_atoms, elem_off);
    if (rc) return rc;
    FusedWorkspace ws = fused_layout(b, batch, n_atoms, mode == 0, (char*)workspace, true);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "simulate_water: workspace needs %zu bytes", ws.bytes);
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.mol_off, mol_off_host, sizeof(int64_t) * (batch + 1),
                                   hipMemcpyHostToDevice, s));
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.elem_off, elem_off.data(), sizeof(int64_t) * elem_off.size(),
                                   hipMemcpyHostToDevice, s));
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.seeds, seeds_host, sizeof(uint64_t) * batch,
                                   hipMemcpyHostToDevice, s));
    WaterFill wf;
    wf.on = true;
    wf.idx_h = idx_h;
    wf.idx_o = idx_o;
    // the same float32 constants as emsim_hip_add_water
    wf.lam_o = (float)voxel_water;
    wf.lam_h = (float)(2.0 * voxel_water);
    wf.exp_o = (float)std::exp(-voxel_water);
    wf.exp_h = (float)std::exp(-2.0 * voxel_water);
    return simulate_core(b, w, coords, n_atoms, ws, batch, psi0, cs_mm, defocus, aperture,
                         symmetric_bandlimit, out, slices_out, mode, s, wf);
}

extern "C" size_t emsim_hip_simulate_raw_workspace_size(const emsim_hip_slice_builder* b,
                                                        int batch, int64_t n_atoms, int mode) {
    if (!b || batch < 1 || n_atoms < 0) return 0;
    return fused_layout(b, batch, n_atoms, mode == 0, nullptr, false, true).bytes;
}

extern "C" int emsim_hip_simulate_raw(const emsim_hip_slice_builder* b,
                                      const emsim_hip_wave_propagator* w, const float* coords,
                                      const uint8_t* elem, int64_t n_atoms,
                                      const int64_t* mol_off_host, int batch, float psi0,
                                      float cs_mm, float defocus, float aperture,
                                      int symmetric_bandlimit, float* out, float* slices_out,
                                      int mode, void* workspace, size_t workspace_bytes,
                                      emsim_stream_t stream_) {
    hipStream_t s = (hipStream_t)stream_;
    if (!mol_off_host || (n_atoms > 0 && !elem))
        return fail(EMSIM_HIP_EINVAL, "simulate_raw: null offsets or element indices");
    int rc = check_simulate(b, w, coords, n_atoms, batch, symmetric_bandlimit, out, mode);
    if (rc) return rc;
    if (mol_off_host[0] != 0 || mol_off_host[batch] != n_atoms)
        return fail(EMSIM_HIP_EINVAL, "simulate_raw: molecule offsets must run from 0 to n_atoms");
    int64_t max_atoms = 0;
    for (int m = 0; m < batch; ++m) {
        const int64_t len = mol_off_host[m + 1] - mol_off_host[m];
        if (len < 0) return fail(EMSIM_HIP_EINVAL, "simulate_raw: offsets must not decrease");
        if (len > max_atoms) max_atoms = len;
    }
    FusedWorkspace ws = fused_layout(b, batch, n_atoms, mode == 0, (char*)workspace, false, true);
    if (!workspace || workspace_bytes < ws.bytes)
        return fail(EMSIM_HIP_EWORKSPACE, "simulate_raw: workspace needs %zu bytes", ws.bytes);
    ws.atom_elem = elem;
    ws.max_mol_atoms = max_atoms;
    EMSIM_HIP_CHECK(hipMemcpyAsync(ws.mol_off, mol_off_host, sizeof(int64_t) * (batch + 1),
                                   hipMemcpyHostToDevice, s));
    return simulate_core(b, w, coords, n_atoms, ws, batch, psi0, cs_mm, defocus, aperture,
                         symmetric_bandlimit, out, slices_out, mode, s);
}

namespace {

int simulate_core(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                  const float* coords, int64_t n_atoms, const FusedWorkspace& ws, int batch,
                  float psi0, float cs_mm, float defocus, float aperture, int symmetric_bandlimit,
                  float* out, float* slices_out, int mode, hipStream_t s, const WaterFill& water) {
    int rc = EMSIM_HIP_OK;
    BatchGeom bg;
    bg.g.e0 = b->edges;
    bg.g.e1 = b->edges + (b->n_slices + 1);
    bg.g.e2 = bg.g.e1 + (b->n1 + 1);
    bg.g.n0 = b->n_slices;
    bg.g.n1 = b->n1;
    bg.g.n2 = b->n2;
    bg.g.last0 = b->last_exact[0];
    bg.g.last1 = b->last_exact[1];
    bg.g.last2 = b->last_exact[2];
    bg.g.inv0 = b->inv_d[0];
    bg.g.inv1 = b->inv_d[1];
    bg.g.inv2 = b->inv_d[2];
    bg.g.one_slice = b->one_slice;
    bg.g.tile_rows = 1;
    bg.g.tiles_per_plane = b->n1;
    bg.n_elems = b->n_elems;
    bg.batch = batch;

    // 1. bucket the atoms by (molecule, slice, element, row)
    const int64_t n_scan = ws.n_rows + 1;
    const int64_t n_blocks = ceil_div(n_scan, kScanTile);
    EMSIM_LAUNCH(0, k_fill_u32, dim3(grid_for(n_scan, 256)), dim3(256), 0, s, ws.row_count,
                 n_scan, 0u);
    if (ws.atom_elem && n_atoms > 0) {
        // molecules as they come: their bounding boxes, for centring on the fly
        EMSIM_LAUNCH(0, k_bbox_init, dim3((unsigned)ceil_div(6 * (int64_t)batch, 256)), dim3(256),
                     0, s, ws.bbox, batch);
        const int64_t chunks = ceil_div(ws.max_mol_atoms > 0 ? ws.max_mol_atoms : 1,
                                        256 * kBoxPerThread);
        EMSIM_LAUNCH(0, k_mol_bbox, dim3((unsigned)chunks, (unsigned)batch), dim3(256), 0, s,
                     coords, (const int64_t*)ws.mol_off, ws.atom_elem, b->n_elems, ws.bbox);
    }
    if (n_atoms > 0)
        EMSIM_LAUNCH(0, k_csr_classify, dim3(grid_for(n_atoms, 256)), dim3(256), 0, s, coords,
                     n_atoms, ws.mol_off, ws.elem_off, bg, ws.atom_row, ws.atom_y, ws.row_count,
                     ws.atom_elem, (const uint32_t*)ws.bbox);
    EMSIM_LAUNCH(0, k_scan_tiles, dim3((unsigned)n_blocks), dim3(kScanBlock), 0, s, ws.row_count,
                 ws.row_start, n_scan, ws.block_sums);
    EMSIM_LAUNCH(0, k_scan_sums, dim3(1), dim3(kScanBlock), 0, s, ws.block_sums, n_blocks);
    EMSIM_LAUNCH(0, k_scan_add, dim3((unsigned)n_blocks), dim3(kScanBlock), 0, s, ws.row_start,
                 n_scan, ws.block_sums);
    if (n_atoms > 0) {
        EMSIM_LAUNCH(0, k_csr_scatter, dim3(grid_for(n_atoms, 256)), dim3(256), 0, s, n_atoms,
                     ws.atom_row, ws.atom_y, ws.row_start, ws.row_count, ws.bucket);
        // a fixed order of every row's atoms for the float row sums
        EMSIM_LAUNCH(0, k_csr_sort_rows, dim3((unsigned)ceil_div(ws.n_rows, kSortRows)),
                     dim3(kSortRows), 0, s, ws.n_rows, ws.row_start, ws.bucket);
    }
    // per-slice molecule order of the spectrum pass (identity beyond one
    // block's worth of molecules)
    const int32_t* order = nullptr;
    if (batch > 1 && batch <= 1024) {
        EMSIM_LAUNCH(0, k_slice_order, dim3((unsigned)b->n_slices), dim3((unsigned)batch), 0, s,
                     ws.row_start, b->n_slices, b->n_elems, b->n1, batch, ws.order);
        order = ws.order;
    }

    // 2. the slices
    float2* wave = mode == 0 ? ws.wave : (float2*)out;
    const int ldh = fused_ld(b);
    const int r = (b->n1 < b->n2 ? b->n1 : b->n2) / 2;
    const float inv_npix = 1.0f / (float)((int64_t)b->n1 * b->n2);
    const int64_t npix = (int64_t)b->n1 * b->n2;
    const int64_t rows = (int64_t)batch * b->n1;
    // a one-slice builder is the weak-phase-object path (emsim/pipe.py:57-64):
    // its slice is followed by singleslice_propagate(dz = 0), the mask alone
    const PropagatorMul prop = make_propagator(w, b->one_slice ? 0.f : (float)b->dz);
    // Columns the propagator zeroes whole are left out between the passes:
    // the wave buffer starts as zeros, pass A never stores into them and
    // pass B skips the tiles that hold only such columns (zeros from the
    // others).  The last pass B runs every tile.
    ColDead dead = dead_columns(prop);
    dead.skip_tiles = true;
    ColDead dead_last = dead;
    dead_last.skip_tiles = false;
    if (dead.lo <= dead.hi) {
        const int64_t n_wave = (int64_t)batch * npix;
        EMSIM_LAUNCH(0, k_fill_complex, dim3(grid_for(n_wave, 256)), dim3(256), 0, s, wave,
                     n_wave, 0.f);
    }
    for (int sl = 0; sl < b->n_slices; ++sl) {
        const bool last = sl == b->n_slices - 1;
        if (!water.on) {
            rc = spec_sparse(b->n1, ws.row_start, ws.bucket, b->sf, b->sf_tiles, ws.Y, sl,
                             b->n_slices,
                             b->n_elems, b->H, ldh, b->n2, symmetric_bandlimit ? 1 : 0, r * r,
                             inv_npix, batch, order, b->tw, s);
            if (rc) return rc;
            rc = slice_pass_a(b->n2, sl == 0, ws.Y, wave, rows / 2, ldh, psi0, w->factor,
                              slices_out, sl, b->n_slices, b->n1, dead, b->tw, s);
            if (rc) return rc;
        } else {
            // water: dense count planes of this slice (atoms + water), then
            // the reference-shaped R2C and spectrum column pass; Y = X[0]
            const int ldx = half_ld(b->H);
            const size_t lds = sizeof(float) * (size_t)b->n_elems * b->n2;
            EMSIM_LAUNCH(0, k_dense_slice_rows, dim3((unsigned)(batch * (int64_t)b->n1)),
                         dim3(256), lds, s, ws.row_start, ws.bucket, sl, b->n_slices,
                         b->n_elems, b->n1, b->n2, batch, ws.hist, 1, water.idx_h, water.idx_o,
                         water.lam_o, water.exp_o, water.lam_h, water.exp_h,
                         (const uint64_t*)ws.seeds);
            rc = r2c_pairs(b->n2, ws.hist, ws.X, (int64_t)b->n_elems * batch * (b->n1 / 2), b->n1,
                           batch, batch, 0, ldx, b->tw, s);
            if (rc) return rc;
            rc = col_spectrum(b->n1, ws.X, b->sf, b->n_elems, batch, b->H, ldx, b->n2,
                              symmetric_bandlimit ? 1 : 0, r * r, inv_npix, b->tw, s);
            if (rc) return rc;
            rc = slice_pass_a(b->n2, sl == 0, ws.X, wave, rows / 2, ldx, psi0, w->factor,
                              slices_out, sl, b->n_slices, b->n1, dead, b->tw, s);
            if (rc) return rc;
        }
        if (mode == 0 && last) {
            LensPropMul lp;
            lp.p = prop;
            lp.l = make_lens(w, cs_mm, defocus, aperture);
            rc = col_lensprop(b->n1, wave, b->n2, b->n2, npix, batch, lp, b->tw, s, dead_last);
        } else {
            rc = col_prop(b->n1, wave, b->n2, b->n2, npix, batch, prop, b->tw, s,
                          last ? dead_last : dead);
        }
        if (rc) return rc;
    }
    // 3. back to real space: |psi|^2 of the image wave, or the exit wave
    if (mode == 0) return row_ifft_intensity(b->n2, wave, out, rows, b->tw, s);
    return row_fft(b->n2, true, wave, wave, rows, b->tw, s);
}

}  // namespace
