/*
 * CPU restatement of the reference's fftw backend -- TEST INFRASTRUCTURE.
 * Used only as the multi-core CPU baseline of bench.py and checked against
 * oracle/emsim_oracle.py in tests/test_oracle_port.py; the product never
 * links it.
 *
 * Follows, per image:
 *   binning      emsim/backend/slice_builder_base.py:180-199 (numpy float64
 *                arange edges, searchsorted(side='right') - 1, closed last edge)
 *   potential    emsim/backend/fftw_ext/src/SliceBuilder.cpp:27-85 (R2C of
 *                every (element, slice) plane, float32 sum over elements in
 *                element order, / nPix, C2R), then the band-limit and clip of
 *                slice_builder_fftw.py:43-48 folded into the spectrum (same
 *                result: the disk mask is real, 0/1 and symmetric)
 *   multislice   emsim/backend/fftw_ext/src/WavePropagator.cpp:41-103
 *   lens, |psi|^2 WavePropagator.cpp:105-138, emsim/pipe.py:79-84
 * FFTW is not available in this image; the transforms are an iterative
 * radix-2 complex FFT (float32 data, twiddles from float64); the R2C of a
 * histogram plane is a complex transform of the real plane and the C2R keeps
 * the real part.  OpenMP parallelises over rows, columns and pixels, as the
 * reference's OpenMP loops do.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float re, im; } cf;

static const float PI32 = 3.14159265358979324f;

/* ---- radix-2 FFT of length n (power of two), in place ------------------ */
static void fft_inplace(cf* x, int n, const cf* tw /* n/2 twiddles exp(-2 pi i k/n) */,
                        int inverse) {
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) { cf t = x[i]; x[i] = x[j]; x[j] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int k = 0; k < half; ++k) {
                cf w = tw[k * step];
                if (inverse) w.im = -w.im;
                cf a = x[i + k], b = x[i + k + half];
                cf bw = {b.re * w.re - b.im * w.im, b.re * w.im + b.im * w.re};
                x[i + k].re = a.re + bw.re; x[i + k].im = a.im + bw.im;
                x[i + k + half].re = a.re - bw.re; x[i + k + half].im = a.im - bw.im;
            }
        }
    }
}

static cf* make_twiddles(int n) {
    cf* tw = (cf*)malloc(sizeof(cf) * (n / 2 + 1));
    for (int k = 0; k < n / 2 + 1; ++k) {
        const double a = -2.0 * 3.14159265358979323846 * k / n;
        tw[k].re = (float)cos(a);
        tw[k].im = (float)sin(a);
    }
    return tw;
}

/* 2-D unnormalised transform of an n x n complex plane */
static void fft2(cf* a, int n, const cf* tw, int inverse) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; ++r) fft_inplace(a + (size_t)r * n, n, tw, inverse);
#pragma omp parallel
    {
        cf* col = (cf*)malloc(sizeof(cf) * n);
#pragma omp for schedule(static)
        for (int c = 0; c < n; ++c) {
            for (int r = 0; r < n; ++r) col[r] = a[(size_t)r * n + c];
            fft_inplace(col, n, tw, inverse);
            for (int r = 0; r < n; ++r) a[(size_t)r * n + c] = col[r];
        }
        free(col);
    }
}

/* ---- binning ------------------------------------------------------------ */
static int ref_edges(int n, double d, double* e) {
    const double start = -d * (double)(n / 2);
    const double stop = start + (double)n * d + d;
    const long len = (long)ceil((stop - start) / d);
    if (len != n + 1) return -1;
    e[0] = start;
    const double next = start + d, delta = next - start;
    e[1] = next;
    for (long i = 2; i < len; ++i) e[i] = start + (double)i * delta;
    return 0;
}

static int bin_index(float x, const double* e, int nb) {
    const double xv = (double)x;
    if (!(xv == xv)) return -1;
    if (xv == e[nb]) return nb - 1;
    int lo = 0, hi = nb + 1;            /* count of edges <= x by bisection */
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (e[mid] <= xv) lo = mid + 1; else hi = mid;
    }
    const int b = lo - 1;
    return (b >= 0 && b < nb) ? b : -1;
}

/* ---- the whole per-image path ------------------------------------------ */
/* exit_wave (nullable): (n, n) complex64 as interleaved floats, the wave
 * after the last slice (before the lens). */
int port_simulate_ex(const float* coords, const int64_t* counts, int n_elems, int n_slices, int n,
                     double dz, double d, const float* sf /* (nE, n, n/2+1) */, float lam,
                     float gamma, float cs_mm, float defocus, float aperture, float psi0,
                     int bandlimit, float* image /* (n, n) */, float* exit_wave) {
    const int H = n / 2 + 1;
    const size_t npix = (size_t)n * n;
    double* e0 = (double*)malloc(sizeof(double) * (n_slices + 1));
    double* e1 = (double*)malloc(sizeof(double) * (n + 1));
    if (ref_edges(n_slices, dz, e0) || ref_edges(n, d, e1)) { free(e0); free(e1); return -1; }
    cf* tw = make_twiddles(n);
    float* hist = (float*)calloc((size_t)n_elems * npix, sizeof(float));
    cf* spec = (cf*)malloc(sizeof(cf) * npix);
    cf* plane = (cf*)malloc(sizeof(cf) * npix);
    cf* wave = (cf*)malloc(sizeof(cf) * npix);
    float* V = (float*)malloc(sizeof(float) * npix);
    for (size_t i = 0; i < npix; ++i) { wave[i].re = psi0; wave[i].im = 0.f; }

    /* float32 constants in the reference's order */
    const float dfx = 1.0f / (float)d / (float)n;
    const float fmax = 0.5f / (float)d;
    const float filt = 0.6667f * fmax;
    const float filter2 = filt * filt;
    const float lpd = lam * PI32 * (float)dz;
    const float factor = lam * gamma;
    const float inv_npix = 1.0f / (float)npix;
    const int r = n / 2;

    /* atoms -> (slice, element) once: slice index per atom */
    int64_t n_atoms = 0;
    for (int e = 0; e < n_elems; ++e) n_atoms += counts[e];
    int* a_s = (int*)malloc(sizeof(int) * (n_atoms ? n_atoms : 1));
    int* a_xy = (int*)malloc(sizeof(int) * (n_atoms ? n_atoms : 1));
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < n_atoms; ++a) {
        const int s = bin_index(coords[3 * a + 0], e0, n_slices);
        const int i = bin_index(coords[3 * a + 1], e1, n);
        const int j = bin_index(coords[3 * a + 2], e1, n);
        a_s[a] = (s >= 0 && i >= 0 && j >= 0) ? s : -1;
        a_xy[a] = i * n + j;
    }

    /* atoms grouped by slice (counting sort, atom order kept within a slice) */
    int64_t* s_off = (int64_t*)calloc((size_t)n_slices + 1, sizeof(int64_t));
    int64_t* s_atoms = (int64_t*)malloc(sizeof(int64_t) * (n_atoms ? n_atoms : 1));
    int* a_e = (int*)malloc(sizeof(int) * (n_atoms ? n_atoms : 1));
    {
        int64_t a = 0;
        for (int e = 0; e < n_elems; ++e)
            for (int64_t k = 0; k < counts[e]; ++k, ++a) {
                a_e[a] = e;
                if (a_s[a] >= 0) ++s_off[a_s[a] + 1];
            }
        for (int s = 0; s < n_slices; ++s) s_off[s + 1] += s_off[s];
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n_slices + 1));
        memcpy(fill, s_off, sizeof(int64_t) * ((size_t)n_slices + 1));
        for (a = 0; a < n_atoms; ++a)
            if (a_s[a] >= 0) s_atoms[fill[a_s[a]]++] = a;
        free(fill);
    }

    for (int s = 0; s < n_slices; ++s) {
        /* histogram planes of this slice */
        memset(hist, 0, sizeof(float) * (size_t)n_elems * npix);
        for (int64_t k = s_off[s]; k < s_off[s + 1]; ++k) {
            const int64_t a = s_atoms[k];
            hist[(size_t)a_e[a] * npix + a_xy[a]] += 1.0f;
        }
        /* spectrum: sum_e F_e FFT2(h_e) (float32, element order) */
        memset(spec, 0, sizeof(cf) * npix);
        for (int e = 0; e < n_elems; ++e) {
            const float* h = hist + (size_t)e * npix;
            for (size_t i = 0; i < npix; ++i) { plane[i].re = h[i]; plane[i].im = 0.f; }
            fft2(plane, n, tw, 0);
            const float* f = sf + (size_t)e * n * H;
#pragma omp parallel for schedule(static)
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    /* F(k1, k2) for k2 > n/2 from the half table: F(-k1, n-k2) */
                    const float fe = j < H ? f[(size_t)i * H + j]
                                           : f[(size_t)((n - i) % n) * H + (n - j)];
                    spec[(size_t)i * n + j].re += plane[(size_t)i * n + j].re * fe;
                    spec[(size_t)i * n + j].im += plane[(size_t)i * n + j].im * fe;
                }
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const int ki = i < n / 2 ? i : i - n, kj = j < n / 2 ? j : j - n;
                const float m = (!bandlimit || ki * ki + kj * kj <= r * r) ? inv_npix : 0.f;
                spec[(size_t)i * n + j].re *= m;
                spec[(size_t)i * n + j].im *= m;
            }
        fft2(spec, n, tw, 1);
        for (size_t i = 0; i < npix; ++i) V[i] = spec[i].re > 1e-13f ? spec[i].re : 1e-13f;
        /* transmit */
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < npix; ++i) {
            const float ph = V[i] * factor;
            const float c = cosf(ph), sn = sinf(ph);
            const cf w = wave[i];
            wave[i].re = w.re * c - w.im * sn;
            wave[i].im = w.re * sn + w.im * c;
        }
        /* propagate */
        fft2(wave, n, tw, 0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const float fx = (float)(i < n / 2 ? i : i - n) * dfx;
                const float fy = (float)(j < n / 2 ? j : j - n) * dfx;
                const float f2 = fx * fx + fy * fy;
                cf* w = wave + (size_t)i * n + j;
                if (f2 <= filter2) {
                    const float ph = lpd * f2;
                    const float pr = cosf(ph), pi = -sinf(ph);
                    const cf x = *w;
                    w->re = (x.re * pr - x.im * pi) / (float)npix;
                    w->im = (x.re * pi + x.im * pr) / (float)npix;
                } else {
                    w->re = 0.f;
                    w->im = 0.f;
                }
            }
        fft2(wave, n, tw, 1);
    }
    if (exit_wave) memcpy(exit_wave, wave, sizeof(cf) * npix);
    /* lens and |psi|^2 */
    {
        float c1 = 0.5f * PI32 * cs_mm * 1e7f * lam * lam * lam;
        float c2 = PI32 * defocus * lam;
        const float fa = aperture / lam;
        fft2(wave, n, tw, 0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const float fx = (float)(i < n / 2 ? i : i - n) * dfx;
                const float fy = (float)(j < n / 2 ? j : j - n) * dfx;
                const float f2 = fx * fx + fy * fy;
                cf* w = wave + (size_t)i * n + j;
                if (f2 <= fa * fa) {
                    const float chi = c1 * f2 * f2 - c2 * f2;
                    const float hr = cosf(chi), hi = -sinf(chi);
                    const cf x = *w;
                    w->re = (x.re * hr - x.im * hi) / (float)npix;
                    w->im = (x.re * hi + x.im * hr) / (float)npix;
                } else {
                    w->re = 0.f;
                    w->im = 0.f;
                }
            }
        fft2(wave, n, tw, 1);
        for (size_t i = 0; i < npix; ++i) image[i] = wave[i].re * wave[i].re + wave[i].im * wave[i].im;
    }
    free(a_s); free(a_xy); free(a_e); free(s_off); free(s_atoms); free(e0); free(e1); free(tw);
    free(hist); free(spec); free(plane); free(wave); free(V);
    return 0;
}

int port_simulate(const float* coords, const int64_t* counts, int n_elems, int n_slices, int n,
                  double dz, double d, const float* sf, float lam, float gamma, float cs_mm,
                  float defocus, float aperture, float psi0, int bandlimit, float* image) {
    return port_simulate_ex(coords, counts, n_elems, n_slices, n, dz, d, sf, lam, gamma, cs_mm,
                            defocus, aperture, psi0, bandlimit, image, NULL);
}

int port_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
