/*
 * emsim_hip.h -- C ABI of the MI355X (gfx950) multislice backend.
 *
 * Plain pointers and sizes only: every array argument named *_dev is device
 * memory owned by the caller (the Python layer allocates it with torch's
 * caching allocator); *_host arguments are host memory read during the call.
 * All launches go to the caller's stream; nothing here synchronises the
 * device except the create functions, which upload constant tables.
 *
 * Every function returns 0 on success or a negative EMSIM_HIP_E* code;
 * emsim_hip_last_error() returns the message of the calling thread's most
 * recent failure.
 *
 * Each entry point names the reference interface it replaces (paths are
 * relative to the reference repository Varato/emsim).
 */
#ifndef EMSIM_HIP_H
#define EMSIM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMSIM_HIP_ABI_VERSION 2

#define EMSIM_HIP_OK 0
#define EMSIM_HIP_EINVAL -1     /* bad argument (shape, size, null pointer) */
#define EMSIM_HIP_EUNSUPPORTED -2 /* size outside what the kernels implement */
#define EMSIM_HIP_EWORKSPACE -3 /* workspace smaller than required */
#define EMSIM_HIP_EDEVICE -4    /* HIP runtime error */
#define EMSIM_HIP_ENOMEM -5     /* host or device allocation failed */

typedef struct ihipStream_t* emsim_stream_t; /* == hipStream_t */

/* ---- library ---------------------------------------------------------- */
int emsim_hip_abi_version(void);
const char* emsim_hip_last_error(void);
/* Lateral sizes: 1 for a power of two from 8 to 4096 (native FFT kernels,
 * every entry point), 2 for any other size from 8 to 2048 (chirp-z
 * transforms on the power-of-two kernels: the per-call slice builder and
 * wave propagator entry points; emsim_hip_simulate* refuse them), 0 else.
 * The reference's FFTW plans take any size (fftw_ext/src/WavePropagator.cpp
 * :25-33, SliceBuilder.cpp:41-55). */
int emsim_hip_fft_size_supported(int n);

/* Host-only helper, exported for the binning tests: the float64 bin edges
 * the reference builds with numpy.arange(s, s + n d + d, d), s = -d (n//2)
 * (emsim/backend/slice_builder_base.py:180-186,202-208), and the float32
 * table the kernels compare against (each edge rounded up to float32).
 * Writes at most cap entries; *len_out = number of edges numpy would make. */
int emsim_hip_bin_edges(int n, double d, double* edges64_host, float* edges32_host, int cap,
                        int* len_out);

/* ---- slice builder ---------------------------------------------------- */
/* Replaces the fftw pybind class slice_kernel.MultiSlicesBuilder
 * (emsim/backend/fftw_ext/src/slice_kernel.cpp:20-64, SliceBuilder.cpp:13-85)
 * and the cuda OneSliceBuilder / MultiSlicesBuilder
 * (emsim/backend/cuda_ext/src/slice_kernel_cuda.cpp:179-214).
 * scattering_factors_host: (n_elems, n1, n2/2+1) float32, the table built by
 * emsim/backend/slice_builder_base.py:30-33.  one_slice != 0 builds the 2-D
 * (one-slice) variant: n_slices must be 1 and coordinate column 0 is ignored
 * (slice_builder_base.py:202-225). */
typedef struct emsim_hip_slice_builder emsim_hip_slice_builder;

int emsim_hip_slice_builder_create(emsim_hip_slice_builder** out,
                                   const float* scattering_factors_host, int n_elems,
                                   int n_slices, int n1, int n2, double dz, double d1,
                                   double d2, int one_slice);
int emsim_hip_slice_builder_destroy(emsim_hip_slice_builder* b);

/* Workspace bytes for binning n_atoms atoms. */
size_t emsim_hip_bin_workspace_size(const emsim_hip_slice_builder* b, int64_t n_atoms);

/* Replaces _bin_atoms_multi_slices / _bin_atoms_one_slice
 * (slice_builder_base.py:180-225) and the cuda binAtoms*Kernel launchers
 * (emsim/backend/cuda_ext/src/SliceBuilder_kernel.cu:290-317).
 * coords_dev: (n_atoms, 3) float32 rows (z, x, y), sorted by element;
 * counts_host: (n_elems) atoms per element, sum <= n_atoms.
 * hist_dev: (n_elems, n_slices, n1, n2) float32 integer-valued counts
 * ((n_elems, n1, n2) for a one-slice builder); numpy.histogramdd semantics
 * on the reference's float64 edges, bit-exact. */
int emsim_hip_bin_atoms(const emsim_hip_slice_builder* b, const float* coords_dev,
                        const int64_t* counts_host, int64_t n_atoms, float* hist_dev,
                        void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream);

/* Replaces MultiSlicesBuilderBase.add_water (slice_builder_base.py:134-147)
 * and the cuda variant (emsim/backend/slice_builder_cuda.py:72-85): every
 * voxel whose count is zero for all elements gets Poisson(voxel_water)
 * oxygens and Poisson(2 voxel_water) hydrogens, added in place to the
 * planes of elements idx_o and idx_h.  voxel_water = water_num_dens d1 d2 dz.
 * The reference draws from numpy's unseeded global RNG; here the draws are
 * a pure function of seed and voxel index. */
int emsim_hip_add_water(const emsim_hip_slice_builder* b, float* hist_dev, int idx_h, int idx_o,
                        double voxel_water, uint64_t seed, emsim_stream_t stream);

/* Workspace bytes for make_slices when at most max_chunk slices are
 * transformed per pass (max_chunk <= 0: all slices at once). */
size_t emsim_hip_slices_workspace_size(const emsim_hip_slice_builder* b, int max_chunk);

/* Replaces MultiSlicesBuilder::makeMultiSlices (SliceBuilder.cpp:27-85)
 * followed by the numpy band-limit and clip (slice_builder_fftw.py:43-48),
 * and make_one_slice (slice_builder_fftw.py:24-32).
 * hist_dev: (n_elems, n_slices, n1, n2) float32; out_dev: (n_slices, n1, n2)
 * float32 projected potential, clipped at 1e-13.  symmetric_bandlimit
 * requires n1 == n2 (the reference's mask only broadcasts then). */
int emsim_hip_make_slices(const emsim_hip_slice_builder* b, const float* hist_dev,
                          float* out_dev, int symmetric_bandlimit, void* workspace_dev,
                          size_t workspace_bytes, emsim_stream_t stream);

/* ---- wave propagator -------------------------------------------------- */
/* Replaces the fftw pybind class wave_kernel.WavePropagator
 * (emsim/backend/fftw_ext/src/wave_kernel.cpp:19-114, WavePropagator.cpp)
 * and the cuda one (emsim/backend/cuda_ext/src/wave_kernel_cuda.cpp:28-143).
 * wave_length / relativity_gamma are the float32 values the reference's
 * binding receives.  Every wave argument is (batch, n1, n2) complex64 stored
 * as interleaved float pairs; in-place calls (out == in) are allowed. */
typedef struct emsim_hip_wave_propagator emsim_hip_wave_propagator;

int emsim_hip_wave_propagator_create(emsim_hip_wave_propagator** out, int n1, int n2,
                                     float d1, float d2, float wave_length,
                                     float relativity_gamma);
int emsim_hip_wave_propagator_destroy(emsim_hip_wave_propagator* w);

/* wave_propagator_base.py:25-29: every pixel = (value, 0). */
int emsim_hip_init_wave(const emsim_hip_wave_propagator* w, float value, float* wave_dev,
                        int batch, emsim_stream_t stream);

/* WavePropagator::sliceTransmit (WavePropagator.cpp:41-55).
 * slice_dev: (batch, n1, n2) float32, or (n1, n2) shared when
 * slice_batch_stride == 0. */
int emsim_hip_slice_transmit(const emsim_hip_wave_propagator* w, const float* wave_dev,
                             const float* slice_dev, int64_t slice_batch_stride,
                             float* out_dev, int batch, emsim_stream_t stream);

/* WavePropagator::spacePropagate (WavePropagator.cpp:57-88). */
int emsim_hip_space_propagate(const emsim_hip_wave_propagator* w, const float* wave_dev,
                              float dz, float* out_dev, int batch, emsim_stream_t stream);

/* WavePropagator::singleSlicePropagate / multiSlicePropagate
 * (WavePropagator.cpp:90-103).  slices_dev: (batch, n_slices, n1, n2)
 * float32, or one (n_slices, n1, n2) stack shared by the batch when
 * slices_batch_stride == 0. */
int emsim_hip_multislice_propagate(const emsim_hip_wave_propagator* w, const float* wave_dev,
                                   const float* slices_dev, int n_slices,
                                   int64_t slices_batch_stride, float dz, float* out_dev,
                                   int batch, emsim_stream_t stream);

/* WavePropagator::lensPropagate (WavePropagator.cpp:105-138). */
int emsim_hip_lens_propagate(const emsim_hip_wave_propagator* w, const float* wave_dev,
                             float cs_mm, float defocus, float aperture, float* out_dev,
                             int batch, emsim_stream_t stream);

/* lensPropagate followed by |psi|^2 (emsim/pipe.py:66-84), fused: the
 * intensity is written without materialising the image wave.
 * scratch_dev: (batch, n1, n2) complex64; it may be wave_dev itself, which
 * is then overwritten.  image_dev: (batch, n1, n2) float32. */
int emsim_hip_lens_intensity(const emsim_hip_wave_propagator* w, const float* wave_dev,
                             float cs_mm, float defocus, float aperture, float* scratch_dev,
                             float* image_dev, int batch, emsim_stream_t stream);

/* |psi|^2 of n complex values (emsim/pipe.py:76,82). */
int emsim_hip_intensity(const float* wave_dev, float* out_dev, int64_t n,
                        emsim_stream_t stream);

/* ---- fused batched pipeline --------------------------------------------- */
/* Replaces, for a batch of molecules, the per-image sequence of
 * emsim/pipe.py:46-84 (build_multi_slices -> init_wave ->
 * multislice_propagate -> lens_propagate -> |psi|^2) as driven per molecule
 * by emsim/simulator.py:71-91, with the slice count fixed by the builder
 * (pot.build_multi_slices(..., n_slices=...)).  No dense histogram,
 * per-element spectrum or potential slice is written to memory.
 * coords_dev: (n_atoms, 3) float32 (z, x, y) of all molecules, molecule m's
 * atoms at [mol_off_host[m], mol_off_host[m+1]), sorted by element;
 * elem_counts_host: (batch, n_elems) atoms per element of each molecule, in
 * the builder's element order; psi0: the incident plane-wave amplitude
 * (sqrt(dose d1 d2), wave_propagator_base.py:25-29).
 * mode 0: out_dev = (batch, n1, n2) float32 images (lens and |psi|^2 with
 * cs_mm, defocus, aperture); mode 1: out_dev = (batch, n1, n2) complex64 exit
 * waves (the lens arguments are ignored).  slices_out_dev (nullable)
 * receives the projected potential slices, (batch, n_slices, n1, n2)
 * float32, as make_slices would return them.
 * With a one-slice builder this is the weak-phase-object path of
 * emsim/pipe.py:57-64,73-77 for the whole batch: the 2-D projected potential
 * (z ignored, slice_builder_base.py:202-225), transmission, then
 * singleslice_propagate(dz = 0) -- the propagator's mask alone. */
size_t emsim_hip_simulate_workspace_size(const emsim_hip_slice_builder* b, int batch,
                                         int64_t n_atoms, int mode);
int emsim_hip_simulate(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                       const float* coords_dev, int64_t n_atoms, const int64_t* mol_off_host,
                       const int64_t* elem_counts_host, int batch, float psi0, float cs_mm,
                       float defocus, float aperture, int symmetric_bandlimit, float* out_dev,
                       float* slices_out_dev, int mode, void* workspace_dev,
                       size_t workspace_bytes, emsim_stream_t stream);

/* emsim_hip_simulate with the water fill of the reference's default image
 * (Pipe(add_water=True), emsim/pipe.py:16,48-52 -> MultiSlicesBuilderBase
 * .add_water, slice_builder_base.py:134-147): in every slice, each voxel
 * empty for all elements gets Poisson(voxel_water) oxygens and
 * Poisson(2 voxel_water) hydrogens in the planes of elements idx_o / idx_h.
 * Molecule m draws with seeds_host[m] exactly the values
 * emsim_hip_add_water draws with that seed for the molecule's own
 * (n_slices, n1, n2) histogram.  Slices are formed from dense count planes
 * (the reference's R2C + spectrum pass) rather than the sparse spectrum
 * pass.  Multi-slice builders only; n_elems * n2 <= 16384. */
size_t emsim_hip_simulate_water_workspace_size(const emsim_hip_slice_builder* b, int batch,
                                               int64_t n_atoms, int mode);
int emsim_hip_simulate_water(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                             const float* coords_dev, int64_t n_atoms, const int64_t* mol_off_host,
                             const int64_t* elem_counts_host, int batch, float psi0, float cs_mm,
                             float defocus, float aperture, int symmetric_bandlimit, int idx_h,
                             int idx_o, double voxel_water, const uint64_t* seeds_host,
                             float* out_dev, float* slices_out_dev, int mode, void* workspace_dev,
                             size_t workspace_bytes, emsim_stream_t stream);

/* emsim_hip_simulate for molecules as they come, with the host preparation
 * (emsim/pipe.py:47 atoms.centralize, emsim/pot.py:40-41
 * atoms.sort_elements_and_count, atoms.py:53-86,123-159) done on the device:
 * coords_dev holds each molecule's atoms uncentred and in any order,
 * elem_dev (n_atoms) the index of each atom's element in the builder's
 * element list (n_elems or more: the atom is not simulated).  Each
 * molecule's bounding box is reduced on the device and its atoms are
 * centred on the fly with the float32 shift atoms.centralize applies, so
 * the images equal emsim_hip_simulate's on the host-prepared copy.  No
 * per-element offsets are needed: atoms are grouped by element as they are
 * bucketed. */
size_t emsim_hip_simulate_raw_workspace_size(const emsim_hip_slice_builder* b, int batch,
                                             int64_t n_atoms, int mode);
int emsim_hip_simulate_raw(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                           const float* coords_dev, const uint8_t* elem_dev, int64_t n_atoms,
                           const int64_t* mol_off_host, int batch, float psi0, float cs_mm,
                           float defocus, float aperture, int symmetric_bandlimit, float* out_dev,
                           float* slices_out_dev, int mode, void* workspace_dev,
                           size_t workspace_bytes, emsim_stream_t stream);

/* ---- learned coarse-slice correction ------------------------------------
 * No counterpart in the reference (emsim has no learned component).  A
 * Fourier neural operator (lifting, `layers` spectral layers of `channels`
 * channels keeping `modes` x-frequencies on each side and `modes`
 * y-frequencies, GELU between layers, linear projection to 2 channels) maps
 * a coarse-slice exit wave plus its projected potentials to the correction
 * towards the fine-slice exit wave.  Activations are channels-last bf16;
 * every dense contraction runs on bf16 MFMA with float32 accumulation. */

#define EMSIM_DT_F32 0
#define EMSIM_DT_BF16 1

/* One batched, strided GEMM, C = A B (+ A2 B2) (+ bias, GELU):
 * element (b, i, k) of A is at a + b a_bs + i a_sm + k a_sk (in elements of
 * a_type), likewise B (b, k, j), C (b, i, j).  k2 > 0 adds a second operand
 * pair summed over k2.  split_k > 0: the k range is cut into
 * S = ceil(k / split_k) chunks and chunk s of batch entry b writes its
 * float32 partial sum to c + (b S + s) c_split_stride (c_bs is unused; c
 * must hold batch S partials; no bias / activation / beta).  beta != 0 adds
 * beta * C_old.
 * z (nullable) receives the pre-activation in C's layout.  Operands are
 * rounded to bf16 on the way into the MFMA. */
typedef struct emsim_gemm_desc {
    int64_t m, n, k, batch;
    const void* a;
    int64_t a_bs, a_sm, a_sk;
    int32_t a_type;
    const void* b;
    int64_t b_bs, b_sk, b_sn;
    int32_t b_type;
    int64_t k2;
    const void* a2;
    int64_t a2_bs, a2_sm, a2_sk;
    int32_t a2_type;
    const void* b2;
    int64_t b2_bs, b2_sk, b2_sn;
    int32_t b2_type;
    void* c;
    int64_t c_bs, c_sm, c_sn;
    int32_t c_type;
    const float* bias;
    int32_t act; /* 0 identity, 1 GELU (erf form) */
    void* z;
    int32_t z_type;
    float beta;
    int64_t split_k, c_split_stride;
} emsim_gemm_desc;
int emsim_hip_gemm(const emsim_gemm_desc* desc, emsim_stream_t stream);

typedef struct emsim_hip_fno emsim_hip_fno;
/* n: lateral size (power of two, 16..4096); in_channels = 2 + number of
 * potential slices; channels 16, 32 or 64; 1 <= modes <= n / 4; 1..8 layers.
 * Parameters start at zero; load them with emsim_hip_fno_set_params. */
int emsim_hip_fno_create(emsim_hip_fno** out, int n, int in_channels, int channels, int modes,
                         int layers);
void emsim_hip_fno_destroy(emsim_hip_fno* f);
/* Number of float32 parameters, and the offsets of the parameter blocks:
 * lift_w (cin_pad, C), lift_b (C), per layer {w (C, C), b (C),
 * r (2 modes, modes, C, C) complex}, proj_w (C, 2), proj_b (2);
 * cin_pad = in_channels rounded up to 16.  cap >= 4 + 3 layers. */
int64_t emsim_hip_fno_param_count(const emsim_hip_fno* f);
int emsim_hip_fno_param_offsets(const emsim_hip_fno* f, int64_t* offsets_host, int cap);
int emsim_hip_fno_set_params(emsim_hip_fno* f, const float* params_dev, emsim_stream_t stream);
int emsim_hip_fno_get_params(const emsim_hip_fno* f, float* params_dev, emsim_stream_t stream);
int emsim_hip_fno_get_grads(const emsim_hip_fno* f, float* grads_dev, emsim_stream_t stream);
size_t emsim_hip_fno_workspace_size(const emsim_hip_fno* f, int batch, int training);
/* x_dev (batch, n, n, cin_pad) bf16 = [Re psi / psi0, Im psi / psi0,
 * sigma V_1 .. sigma V_ns, 0...] from wave_dev (batch, n, n) complex64 and
 * slices_dev (batch, ns, n, n) float32. */
int emsim_hip_fno_pack_input(const emsim_hip_fno* f, const float* wave_dev,
                             const float* slices_dev, int n_slices, float psi0, float sigma,
                             int batch, void* x_dev, emsim_stream_t stream);
/* out_dev (batch, n, n, 2) float32.  training != 0 keeps what the backward
 * pass needs in the workspace (sized with training = 1). */
int emsim_hip_fno_forward(emsim_hip_fno* f, const void* x_dev, float* out_dev, int batch,
                          int training, void* workspace_dev, size_t workspace_bytes,
                          emsim_stream_t stream);
/* Inference straight from the simulation: emsim_hip_fno_pack_input fused
 * into the lifting (the same values, no packed tensor in memory).  Same
 * workspace as an inference forward pass.  Returns EMSIM_HIP_EUNSUPPORTED
 * for network shapes the fused row kernels do not serve (use pack_input +
 * forward then). */
int emsim_hip_fno_infer(emsim_hip_fno* f, const float* wave_dev, const float* slices_dev,
                        int n_slices, float psi0, float sigma, float* out_dev, int batch,
                        void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream);
/* on != 0: every forward pass uses the generic GEMM sequence instead of the
 * fused row kernels (which serve 32 channels, <= 16 modes, n <= 512); the
 * two compute the same network, this selects one for cross-checks. */
int emsim_hip_fno_force_generic(int on);
/* Gradients of the parameters (overwritten) for dout_dev = dL / d out, after
 * a training forward pass with the same x_dev and workspace. */
int emsim_hip_fno_backward(emsim_hip_fno* f, const void* x_dev, const float* dout_dev, int batch,
                           void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream);
int emsim_hip_fno_adam(emsim_hip_fno* f, float lr, float beta1, float beta2, float eps,
                       float weight_decay, int step, emsim_stream_t stream);
/* wave_dev = coarse_wave_dev + psi0 (out_r + i out_i), complex64. */
int emsim_hip_fno_apply(const emsim_hip_fno* f, const float* coarse_wave_dev, const float* out_dev,
                        float psi0, float* wave_dev, int batch, emsim_stream_t stream);
/* loss_dev (batch, 3): |pred - target|^2 / |target|^2, |target|^2 and
 * |pred|^2 per image (complex64 waves of npix pixels); dout_dev (nullable,
 * batch npix 2):
 * the gradient of the batch-mean loss with respect to the network output
 * when pred = coarse + psi0 out. */
size_t emsim_hip_fno_loss_workspace_size(int64_t npix, int batch);
int emsim_hip_fno_loss(const float* pred_dev, const float* target_dev, int64_t npix, int batch,
                       float psi0, float* loss_dev, float* dout_dev, void* workspace_dev,
                       size_t workspace_bytes, emsim_stream_t stream);

/* ---- one molecule in many orientations ---------------------------------
 * The device-side counterpart of emsim/atoms.py:162-211 (rotate /
 * orientations_gen) followed by atoms.centralize (:123-159): instance b of
 * out_dev (batch, n_atoms, 3) is rot_dev[b] (row-major 3x3 float32) applied to
 * (x - c0) for every atom x of base_dev (n_atoms, 3), then translated so its
 * bounding box is centred on the origin (float32, the host's operation
 * order).  Atom order is kept, so a base sorted by element stays sorted. */
size_t emsim_hip_orient_workspace_size(int batch);
int emsim_hip_orient_batch(const float* base_dev, int64_t n_atoms, const float* rot_dev,
                           float c0x, float c0y, float c0z, int batch, float* out_dev,
                           void* workspace_dev, size_t workspace_bytes, emsim_stream_t stream);

/* emsim_hip_orient_batch without a workspace argument; it allocates nothing.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops).  The
 * bounding-box keys live in a library-owned device array (96 KiB on each GPU
 * the library runs on, held until the process ends), so batch is at most
 * EMSIM_HIP_ORIENT_NOALLOC_MAX_BATCH (larger batches: EMSIM_HIP_EUNSUPPORTED).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (concurrent streams, several host threads) race and must be serialised
 * by the caller.  The first call on a device looks up the array's address
 * (hipGetSymbolAddress), so make it once outside graph capture. */
#define EMSIM_HIP_ORIENT_NOALLOC_MAX_BATCH 4096
int emsim_hip_orient_batch_noalloc(const float* base_dev, int64_t n_atoms, const float* rot_dev,
                                   float c0x, float c0y, float c0z, int batch, float* out_dev,
                                   emsim_stream_t stream);

/* ---- launch-shape knobs (no counterpart in the reference; set by the
 * autotuner, emsim_amd/autotune.py) --------------------------------------
 * "row_items_per_block": rows (row pairs) each row-worker block handles,
 * 1..64; "col_tiles_per_block": column tiles each column-kernel block
 * handles, 1..32; 0 restores the built-in rule.  Process-wide; they change
 * how work is split over blocks, never the results. */
int emsim_hip_tuning_set(const char* key, int value);
/* The current value (0 = built-in rule), or a negative status. */
int emsim_hip_tuning_get(const char* key);

/* emsim_hip_simulate with the offsets already in device memory:
 * mol_off_dev (batch + 1) and elem_off_dev (batch, n_elems + 1) absolute atom
 * offsets of each molecule and of each molecule's element groups (the
 * caller's responsibility to make consistent).  It reads no host memory and
 * allocates nothing, so a stream capturing a HIP graph can record it (after
 * one ordinary call has set up the library's per-device tables). */
int emsim_hip_simulate_dev(const emsim_hip_slice_builder* b, const emsim_hip_wave_propagator* w,
                           const float* coords_dev, int64_t n_atoms, const int64_t* mol_off_dev,
                           const int64_t* elem_off_dev, int batch, float psi0, float cs_mm,
                           float defocus, float aperture, int symmetric_bandlimit, float* out_dev,
                           float* slices_out_dev, int mode, void* workspace_dev,
                           size_t workspace_bytes, emsim_stream_t stream);

/* ---- instrumentation (no counterpart in the reference, which has no
 * profiler; simulator.py:77,89-91 only logs wall time per image) ---------- */
/* While enabled, every kernel launch is bracketed by two HIP events on its
 * stream. */
int emsim_hip_timing_enable(int on);
/* Waits for the recorded events and writes one line per (kernel, length):
 * "name\tn\tcount\ttotal_ms\tmin_ms\tmax_ms\n"; then forgets the records.
 * *needed receives the buffer size the full text needs. */
int emsim_hip_timing_collect(char* buf, size_t cap, size_t* needed);

#ifdef __cplusplus
}
#endif

#endif /* EMSIM_HIP_H */
