/* hpxmi: MI355X-native HEALPix spherical-harmonic transform and pixel
 * remaps, exported as a plain C ABI.
 *
 * All device pointers are HIP device memory owned by the caller; outputs are
 * allocated by the caller.  Every call is asynchronous on the given stream
 * and returns HPX_OK or an error code; hpx_last_error() then describes the
 * failure (thread-local).  No call allocates device memory except plan
 * creation, hpx_sht_autotune (temporary buffers, freed before it returns)
 * and hpx_sht_mixed_create (the graph it instantiates).
 *
 * Layouts (row-major, last index fastest):
 *   map  f    float32 [nbatch, npix]                 (RING order for HEALPix)
 *   alm  a    complex64 [nbatch, lmax, mmax]         (interleaved re, im;
 *                                                      a[l, m] = 0 for m > l)
 */
#ifndef HPXMI_H
#define HPXMI_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPX_OK 0
#define HPX_ERR_ARG 1          /* invalid argument (size, nside, origin, ...) */
#define HPX_ERR_HIP 2          /* HIP runtime error */
#define HPX_ERR_UNSUPPORTED 3  /* valid request outside this build's limits */

/* Description of the last error on the calling thread. */
const char* hpx_last_error(void);

/* Library ABI version (major * 100 + minor). */
int hpx_version(void);

/* ------------------------------------------------------------------ */
/* Pixel remaps                                                        */
/* ------------------------------------------------------------------ */

/* Remap kinds.  Each replaces one pybind11 export of the reference's
 * cuhpx_remap module (src/data_remapping/hpx_remapping.cpp:452-478); the
 * 1-D and _batch forms are the same call with nrows = 1 or the product of
 * the leading dimensions:
 *   HPX_RING2NEST  ring2nest / ring2nest_batch   (hpx_remapping.cpp:21-32, 186-228)
 *   HPX_NEST2RING  nest2ring / nest2ring_batch   (hpx_remapping.cpp:35-46)
 *   HPX_NEST2FLAT  nest2xy / nest2xy_batch       (hpx_remapping.cpp:48-81)
 *   HPX_FLAT2NEST  xy2nest / xy2nest_batch       (hpx_remapping.cpp:83-116)
 *   HPX_RING2FLAT  ring2xy / ring2xy_batch       (hpx_remapping.cpp:118-150)
 *   HPX_FLAT2RING  xy2ring / xy2ring_batch       (hpx_remapping.cpp:152-184)
 *   HPX_FLAT2FLAT  xy2xy / xy2xy_batch           (hpx_remapping.cpp:406-448)
 */
enum {
  HPX_RING2NEST = 0,
  HPX_NEST2RING = 1,
  HPX_NEST2FLAT = 2,
  HPX_FLAT2NEST = 3,
  HPX_RING2FLAT = 4,
  HPX_FLAT2RING = 5,
  HPX_FLAT2FLAT = 6
};

/* Origin codes of the flat (XY) layout, as hpx_remapping.cpp:65-75. */
enum { HPX_ORIGIN_S = 0, HPX_ORIGIN_E = 1, HPX_ORIGIN_N = 2, HPX_ORIGIN_W = 3 };

/* out[row, sigma(i)] = in[row, i] for every row (a pure permutation of
 * elem_bytes-sized elements, bit-exact for any dtype).  nside must be a
 * power of two; each row holds 12*nside^2 elements; in and out must not
 * overlap.  For the one-sided flat kinds (origin, clockwise) describe the
 * flat layout and the dst_* arguments are ignored; for HPX_FLAT2FLAT they
 * describe the source layout and dst_* the destination layout.
 * elem_bytes is 1, 2, 4 or 8. */
int hpx_remap(int kind, int64_t nside, int origin, int clockwise,
              int dst_origin, int dst_clockwise, const void* in, void* out,
              int64_t nrows, int elem_bytes, hipStream_t stream);

/* Plain device-to-device copy kernel (16-byte vector loads/stores); the
 * bandwidth yardstick that remap and FFT traffic are scaled by.  No
 * counterpart in the reference. */
int hpx_copy(const void* in, void* out, int64_t nbytes, hipStream_t stream);

/* ------------------------------------------------------------------ */
/* Spherical harmonic transform                                        */
/* ------------------------------------------------------------------ */

typedef struct hpx_sht_plan hpx_sht_plan;

/* Legendre-step arithmetic. */
enum {
  HPX_PREC_FP32 = 0,    /* FP32 MFMA, FP32 operands */
  HPX_PREC_BF16X3 = 1,  /* BF16 MFMA, 3-term split (hi*hi + hi*lo + lo*hi), FP32 accumulate */
  HPX_PREC_BF16 = 2     /* BF16 MFMA, single term */
};

/* Adjoint variants. */
enum {
  HPX_ADJ_EXACT = 0,   /* exact adjoint of the forward operator */
  HPX_ADJ_COMPAT = 1   /* the reference's hand-written backward formulas
                          (hpx_sht.py:564-583, 613-641) */
};

/* Plan for a HEALPix grid: the state SHTCUDA / iSHTCUDA build in their
 * constructors (cuhpx/hpx_sht.py:645-683, 696-730): ring geometry, ring
 * quadrature weights (ring_weights[t], t = 0..4*nside-2, FP64, as
 * sht_tools.healpix_weights returns them), and the normalised Legendre table,
 * which is generated on the device in FP64 by three-term recurrence with
 * underflow scaling (replacing sht_tools.py:232-264).  Replaces the plan
 * state of HealpixFFT / HealpixIFFT (src/harmonic_transform/hpx_fft.h:69-255).
 * nside up to 8192 (larger: HPX_ERR_UNSUPPORTED).
 * Synchronises the stream before returning. */
int hpx_sht_plan_create_healpix(int nside, int lmax, int mmax,
                                const double* ring_weights,
                                hipStream_t stream, hpx_sht_plan** plan);

/* Plan for a general iso-latitude grid that is mirror-symmetric about the
 * equator (e.g. an equiangular lat/lon grid for Regridding, hpx_regrid.py:68-86):
 * ring t has nphi[t] pixels at longitudes phi0[t] + 2*pi*j/nphi[t] and
 * colatitude theta[t] (north to south); pixels are stored ring after ring.
 * No counterpart in the reference (which delegates lat/lon grids to
 * torch_harmonics). */
int hpx_sht_plan_create_rings(int nring, const int32_t* nphi,
                              const double* theta, const double* phi0,
                              const double* weights, int lmax, int mmax,
                              hipStream_t stream, hpx_sht_plan** plan);

void hpx_sht_plan_destroy(hpx_sht_plan* plan);

/* Pixels per map and device bytes of the plan's tables. */
int64_t hpx_sht_plan_npix(const hpx_sht_plan* plan);
int64_t hpx_sht_plan_table_bytes(const hpx_sht_plan* plan);

/* Device workspace (bytes) that forward / inverse / adjoint calls on nbatch
 * maps need; the caller allocates it (256-byte aligned). */
int64_t hpx_sht_workspace_bytes(const hpx_sht_plan* plan, int64_t nbatch);

/* Largest nbatch that one call on this plan accepts (maps for a scalar plan,
 * fields for a vector plan): the Legendre GEMMs address one order's data
 * through 32-bit offsets, max(2 * ring pairs, lmax + 32) rows of 2 * nbatch
 * floats.  Larger calls return HPX_ERR_UNSUPPORTED; split the batch.  No
 * counterpart in the reference. */
int64_t hpx_sht_max_batch(const hpx_sht_plan* plan);

/* Forward SHT, f[nbatch, npix] -> alm[nbatch, lmax, mmax] (complex64).
 * Replaces SHTFunction.forward (hpx_sht.py:517-562): cuhpx_fft
 * healpix_rfft_batch / healpix_rfft_class (hpx_fft.cpp:21-84, 152-193)
 * followed by the einsum '...kmn,mlk->...lmn' (hpx_sht.py:557). */
int hpx_sht_forward(const hpx_sht_plan* plan, int precision, const float* f,
                    float* alm, int64_t nbatch, void* workspace,
                    int64_t workspace_bytes, hipStream_t stream);

/* hpx_sht_forward without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a fixed 128 MiB device array of the library (one copy per
 * device, held until the process ends); calls that need more
 * (hpx_sht_workspace_bytes(plan, nbatch) > 134217728) return
 * HPX_ERR_UNSUPPORTED before touching the device, e.g. nside 256 with
 * lmax = mmax = 513 takes at most 16 maps per call.  Makes no allocation;
 * calls that overlap on one device race on the array (serialise them, e.g.
 * on one stream).  Results equal hpx_sht_forward's bit for bit. */
int hpx_sht_forward_noalloc(const hpx_sht_plan* plan, int precision, const float* f,
                            float* alm, int64_t nbatch, hipStream_t stream);

/* Inverse SHT, alm[nbatch, lmax, mmax] -> f[nbatch, npix].
 * Replaces iSHTFunction.forward (hpx_sht.py:588-611): the einsum
 * '...lmn,mlk->...kmn' (hpx_sht.py:599) followed by cuhpx_fft
 * healpix_irfft_batch / healpix_irfft_class (hpx_fft.cpp:86-149, 195-238).
 * The input is not modified (the reference's irfft mutates its input). */
int hpx_sht_inverse(const hpx_sht_plan* plan, int precision, const float* alm,
                    float* f, int64_t nbatch, void* workspace,
                    int64_t workspace_bytes, hipStream_t stream);

/* Gradient of the forward SHT: galm[nbatch, lmax, mmax] -> gf[nbatch, npix].
 * Replaces SHTFunction.backward (hpx_sht.py:564-583). */
int hpx_sht_forward_adjoint(const hpx_sht_plan* plan, int precision, int mode,
                            const float* galm, float* gf, int64_t nbatch,
                            void* workspace, int64_t workspace_bytes,
                            hipStream_t stream);

/* Gradient of the inverse SHT: gf[nbatch, npix] -> galm[nbatch, lmax, mmax].
 * Replaces iSHTFunction.backward (hpx_sht.py:613-641). */
int hpx_sht_inverse_adjoint(const hpx_sht_plan* plan, int precision, int mode,
                            const float* gf, float* galm, int64_t nbatch,
                            void* workspace, int64_t workspace_bytes,
                            hipStream_t stream);

/* ------------------------------------------------------------------ */
/* FP64 transforms (float64 maps, complex128 coefficients)             */
/* ------------------------------------------------------------------ */

/* The reference computes in FP64 when given float64 maps: its FFT dispatch
 * returns complex128 spectra for float64 input (src/harmonic_transform/
 * hpx_fft.cpp:45) and the einsums cast the FP64 table to the input's dtype
 * (cuhpx/hpx_sht.py:557, 599).  These calls are that path: FP64 ring FFTs
 * (Bluestein tables in FP64) and FP64 Legendre GEMMs on an FP64 table.
 *
 * hpx_sht_plan_enable_f64 builds a plan's FP64 state (the FP64 Legendre
 * table in HBM: 8 bytes per (m, l, ring pair) of the folded triangle, 0.6 GB
 * at nside 256, lmax = mmax = 513, twice that for the two tables of a vector
 * plan; Bluestein tables, twiddles, phases); it allocates, synchronises the
 * stream, and is a no-op on a plan that has it.  The _f64 calls below return
 * HPX_ERR_ARG on a plan without it.  FP64 FFTs stop at 8192 points, so
 * longer rings run ring by ring at half or quarter length (nside 2048 and
 * 4096); plans that would still need more (nside 8192) are refused. */
int hpx_sht_plan_enable_f64(hpx_sht_plan* plan, hipStream_t stream);

/* Workspace bytes of an FP64 call on nbatch maps. */
int64_t hpx_sht_workspace_bytes_f64(const hpx_sht_plan* plan, int64_t nbatch);

/* hpx_sht_forward / hpx_sht_inverse / the adjoints in FP64: f float64
 * [nbatch, npix], alm complex128 [nbatch, lmax, mmax] (interleaved re, im). */
int hpx_sht_forward_f64(const hpx_sht_plan* plan, const double* f, double* alm,
                        int64_t nbatch, void* workspace, int64_t workspace_bytes,
                        hipStream_t stream);
int hpx_sht_inverse_f64(const hpx_sht_plan* plan, const double* alm, double* f,
                        int64_t nbatch, void* workspace, int64_t workspace_bytes,
                        hipStream_t stream);
int hpx_sht_forward_adjoint_f64(const hpx_sht_plan* plan, int mode, const double* galm,
                                double* gf, int64_t nbatch, void* workspace,
                                int64_t workspace_bytes, hipStream_t stream);
int hpx_sht_inverse_adjoint_f64(const hpx_sht_plan* plan, int mode, const double* gf,
                                double* galm, int64_t nbatch, void* workspace,
                                int64_t workspace_bytes, hipStream_t stream);

/* hpx_ring_rfft / hpx_ring_irfft in FP64 (complex128 ftm[nbatch, nring, mmax]);
 * they replace cuhpx_fft.healpix_rfft_* / healpix_irfft_* on float64 /
 * complex128 tensors (hpx_fft.cpp:45, 122). */
int hpx_ring_rfft_f64(const hpx_sht_plan* plan, const double* f, double* ftm, int64_t nbatch,
                      hipStream_t stream);
int hpx_ring_irfft_f64(const hpx_sht_plan* plan, const double* ftm, double* f, int64_t nbatch,
                       hipStream_t stream);

/* hpx_vsht_forward / hpx_vsht_inverse in FP64 on a vector plan (after
 * hpx_sht_plan_enable_f64): f float64 [nbatch, 2, npix], alm complex128
 * [nbatch, 2, lmax, mmax]; the reference's VectorSHT / VectoriSHT on float64
 * input (their einsums cast the tables to the input's dtype, hpx_sht.py:
 * 351-372, 413-432).  Workspace: hpx_vsht_workspace_bytes_f64. */
int64_t hpx_vsht_workspace_bytes_f64(const hpx_sht_plan* plan, int64_t nbatch);
int hpx_vsht_forward_f64(const hpx_sht_plan* plan, const double* f, double* alm,
                         int64_t nbatch, void* workspace, int64_t workspace_bytes,
                         hipStream_t stream);
int hpx_vsht_inverse_f64(const hpx_sht_plan* plan, const double* alm, double* f,
                         int64_t nbatch, void* workspace, int64_t workspace_bytes,
                         hipStream_t stream);

/* ------------------------------------------------------------------ */
/* Vector (spheroidal / toroidal) SHT                                  */
/* ------------------------------------------------------------------ */

/* Normalisations of the vector tables (the reference's norm argument,
 * cuhpx/sht_tools.py:194-221): "ortho", "schmidt", and any other string,
 * which the reference scales like HPX_NORM_FOUR_PI. */
enum { HPX_NORM_ORTHO = 0, HPX_NORM_FOUR_PI = 1, HPX_NORM_SCHMIDT = 2 };

/* Plan for the vector transform on a HEALPix grid: the state VectorSHT /
 * VectoriSHT build in their constructors (cuhpx/hpx_sht.py:288-330, 379-402):
 * ring geometry and weights as hpx_sht_plan_create_healpix, and instead of
 * the scalar table the two tables of _precompute_dlegpoly
 * (cuhpx/sht_tools.py:267-304), D0 = dP/dtheta and D1 = m P / sin(theta),
 * generated on the device in FP64 (Condon-Shortley sign on odd m when
 * csphase != 0, the reference's default).  norm scales the P rows the tables
 * are built from as legpoly(norm=..., inverse=...) does (sht_tools.py:
 * 196-221): by sqrt(4 pi) unless HPX_NORM_ORTHO, and for HPX_NORM_SCHMIDT by
 * (2 L + 1)^(-1/2) at the row's degree L (l for D0, l + 1 for D1); with
 * inverse != 0 (VectoriSHT's tables, hpx_sht.py:396) by the reciprocals.
 * With HPX_NORM_ORTHO the tables do not depend on inverse.  The scalar
 * entry points reject a vector plan and the vector ones a scalar plan.
 * Synchronises the stream. */
int hpx_vsht_plan_create_healpix(int nside, int lmax, int mmax,
                                 const double* ring_weights, int csphase,
                                 int norm, int inverse,
                                 hipStream_t stream, hpx_sht_plan** plan);

/* Workspace (bytes) for nbatch vector fields (that of 2*nbatch scalar maps). */
int64_t hpx_vsht_workspace_bytes(const hpx_sht_plan* plan, int64_t nbatch);

/* Forward vector SHT, f[nbatch, 2, npix] (theta and phi components) ->
 * alm[nbatch, 2, lmax, mmax] (complex64; spheroidal, toroidal).
 * Replaces VectorSHT.forward (hpx_sht.py:333-374): cuhpx_fft
 * healpix_rfft_batch of both components, then the four real einsums per
 * output part with weights dpct * w / (l (l + 1)). */
int hpx_vsht_forward(const hpx_sht_plan* plan, int precision, const float* f,
                     float* alm, int64_t nbatch, void* workspace,
                     int64_t workspace_bytes, hipStream_t stream);

/* Inverse vector SHT, alm[nbatch, 2, lmax, mmax] -> f[nbatch, 2, npix].
 * Replaces VectoriSHT.forward (hpx_sht.py:404-444): the four real einsums
 * per component, then cuhpx_fft healpix_irfft_batch.  The input is not
 * modified. */
int hpx_vsht_inverse(const hpx_sht_plan* plan, int precision, const float* alm,
                     float* f, int64_t nbatch, void* workspace,
                     int64_t workspace_bytes, hipStream_t stream);

/* Per-ring real FFTs alone, in the reference's ring-major layout
 * ftm[nbatch, nring, mmax] (complex64), where mmax is the plan's mmax
 * (the reference's L).  A plan built with lmax = 0 has no Legendre table and
 * serves only these two calls.
 * hpx_ring_rfft replaces cuhpx_fft.healpix_rfft_batch / healpix_rfft_class
 * (src/harmonic_transform/hpx_fft.cpp:21-84, 152-193):
 *   ftm[b, t, m] = e^{-i m phi0_t} sum_j f[b, off_t + j] e^{-2 pi i m j / N_t}
 *   for m < min(mmax, N_t/2 + 1), 0 above.
 * hpx_ring_irfft replaces cuhpx_fft.healpix_irfft_batch / healpix_irfft_class
 * (hpx_fft.cpp:86-149, 195-238): per ring, irfft(ftm[b, t, :] e^{i m phi0_t},
 * n = N_t, norm = "forward"); unlike the reference it leaves ftm unchanged. */
int hpx_ring_rfft(const hpx_sht_plan* plan, const float* f, float* ftm, int64_t nbatch,
                  hipStream_t stream);
int hpx_ring_irfft(const hpx_sht_plan* plan, const float* ftm, float* f, int64_t nbatch,
                   hipStream_t stream);

/* Split-phase entry points: the two stages of each direction, with the
 * ring-pair spectra left at the start of the workspace between them (layout
 * [mmax][2][kpad][ncol] float32, kpad = ring pairs rounded up to 32,
 * ncol = 2*nbatch rounded up to 32; parity 0 holds S = w (F_t + F_t'),
 * parity 1 holds D = w (F_t - F_t')); the Legendre stages also use the rest
 * of the workspace for the coefficients in GEMM order.  hpx_sht_forward is
 * hpx_sht_stage_ring_analysis followed by hpx_sht_stage_legendre_analysis;
 * hpx_sht_inverse is hpx_sht_stage_legendre_synthesis followed by
 * hpx_sht_stage_ring_synthesis.  A ring pair's spectrum stops at
 * K = min(mmax, nphi / 2 + 1).  The Legendre analysis skips, per m, the
 * leading ring pairs whose spectra stop below m, and reads the others
 * (zeros included) only as far as their spectra are written: the ring
 * analysis writes each pair for every m that the Legendre analysis reads
 * (zeros from K on) and leaves the rest unwritten.  The Legendre synthesis
 * may leave entries with m >= K unwritten, and the ring synthesis never
 * reads them.  No
 * counterpart in the reference (its stages are the cuhpx_fft calls and the
 * einsums inside one autograd Function); they let callers time, overlap or
 * pipeline the stages. */
int hpx_sht_stage_ring_analysis(const hpx_sht_plan* plan, const float* f, int64_t nbatch,
                                void* workspace, int64_t workspace_bytes, hipStream_t stream);
int hpx_sht_stage_legendre_analysis(const hpx_sht_plan* plan, int precision, float* alm,
                                    int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                    hipStream_t stream);
int hpx_sht_stage_legendre_synthesis(const hpx_sht_plan* plan, int precision, const float* alm,
                                     int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                     hipStream_t stream);
int hpx_sht_stage_ring_synthesis(const hpx_sht_plan* plan, float* f, int64_t nbatch,
                                 void* workspace, int64_t workspace_bytes, hipStream_t stream);

/* ------------------------------------------------------------------ */
/* Launch-shape autotuning and mixed-size batched requests             */
/* ------------------------------------------------------------------ */

/* Measures the launch shapes of the ring FFT stage (fields per work-group of
 * each FFT size class, both directions) and of the two Legendre GEMMs
 * (column tiles per wavefront) for this plan, precision and batch on the
 * stream's device, and keeps the fastest in the plan; later calls with the
 * same precision and batch column count use them.  On a vector plan nbatch
 * counts scalar maps (2 per vector field).  Synchronises the stream;
 * must not be called while the stream is being captured.  No counterpart in
 * the reference (its launch shapes are fixed, hpx_fft_cuda.cu:1062-1332). */
int hpx_sht_autotune(hpx_sht_plan* plan, int precision, int64_t nbatch, void* workspace,
                     int64_t workspace_bytes, hipStream_t stream);

/* The launch shapes used for (precision, nbatch), 15 values: shapes[0..2]
 * analysis fields per work-group of FFT classes 0..2, shapes[3..5] the same
 * for synthesis, shapes[6] / shapes[7] forward / inverse GEMM column tiles
 * (forward: 4 = the 8-wavefront kernel, 128 columns per work-group),
 * shapes[8] inverse GEMM ring-pair rows per work-group in units of 64,
 * shapes[9..14] threads per field (64, 128 or 256) of the six FFT entries 0..5. */
int hpx_sht_tuned_shapes(const hpx_sht_plan* plan, int precision, int64_t nbatch,
                         int32_t* shapes);

/* Sets the launch shapes used for (precision, nbatch) in the layout of
 * hpx_sht_tuned_shapes, without timing anything (a saved tuning result, or
 * one kernel variant chosen on purpose); every value is checked against the
 * kernels that exist and the plan's LDS.  Launch shapes never change
 * results.  No counterpart in the reference. */
int hpx_sht_set_shapes(hpx_sht_plan* plan, int precision, int64_t nbatch, const int32_t* shapes);

/* A mixed-size request: nplans HEALPix (or ring-grid) plans, each with its
 * own batch of maps, run as one HIP graph whose branches (one per plan) run
 * concurrently.  direction HPX_MIXED_FORWARD: maps_in[i] -> alm[i];
 * HPX_MIXED_INVERSE: alm[i] -> maps_out[i]; both: maps_in[i] -> alm[i] ->
 * maps_out[i].  The buffers and the workspace are bound at creation (static
 * buffers): each hpx_sht_mixed_launch reads and writes the same addresses.
 * Replaces the reference's per-size SHTCUDA / iSHTCUDA calls (hpx_sht.py:
 * 685-691, 732-735), whose FFT plan cache is rebuilt whenever nside changes
 * (hpx_fft.cpp:49-58).  Launches that overlap on the same request race. */
typedef struct hpx_sht_mixed hpx_sht_mixed;
#define HPX_MIXED_FORWARD 1
#define HPX_MIXED_INVERSE 2
#define HPX_MIXED_MAX_PLANS 16

/* Workspace bytes for a mixed request (sum of the per-plan workspaces,
 * each rounded up to 256 bytes); -1 on a bad argument. */
int64_t hpx_sht_mixed_workspace_bytes(int nplans, const hpx_sht_plan* const* plans,
                                      const int64_t* nbatch);
int hpx_sht_mixed_create(int nplans, const hpx_sht_plan* const* plans, const int64_t* nbatch,
                         int precision, int direction, const float* const* maps_in,
                         float* const* alm, float* const* maps_out, void* workspace,
                         int64_t workspace_bytes, hpx_sht_mixed** mixed);
int hpx_sht_mixed_launch(hpx_sht_mixed* mixed, hipStream_t stream);
void hpx_sht_mixed_destroy(hpx_sht_mixed* mixed);

#ifdef __cplusplus
}
#endif

#endif /* HPXMI_H */
