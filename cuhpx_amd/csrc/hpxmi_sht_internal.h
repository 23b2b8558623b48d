// Internal plan layout and kernel launchers of the SHT (not part of the C ABI).
#pragma once

#include "hpxmi_common.h"

#include <algorithm>
#include <complex>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace hpxmi {

// Launch shapes of the ring FFT stage for one column count: fields per
// work-group and threads per field for the wave-level size classes 0..2, for
// the analysis [0] and synthesis [1] directions.
struct FftTune {
  int waves[2][3] = {{16, 4, 2}, {16, 4, 2}};       // fields per work-group [direction][class]
  int tpf[2][3] = {{64, 128, 64}, {64, 128, 64}};   // threads per field (64, 128 or 256)
};
// Launch shapes of the Legendre GEMMs for one (precision, column count):
// 16-column MFMA tiles per wavefront (work-group column block = 64 * ct).
struct GemmTune {
  int fwd_ct = 1;  // 1: what the tuner picks on MI355X at 64 fields (4 waves per SIMD)
  int inv_ct = 1;
  int inv_rh = 1;  // inverse: 64-ring-pair row halves per work-group (1 or 2)
};
// Allowed values (what the kernels are instantiated for).
constexpr int kFftWaveChoices[3][3] = {{4, 8, 16}, {2, 4, 8}, {1, 2, 4}};
// Threads per field of the FFT kernels: one wavefront, or two or four sharing
// each pass (a work-group barrier between passes); fields * threads <= 1024.
constexpr int kFftTpfChoices[3] = {64, 128, 256};
constexpr int kGemmCtChoices[2] = {1, 2};
// Forward GEMM "column tiles" value that selects the 8-wave kernel (128
// columns per work-group, 64-ring-pair chunks; scalar plans with kpad % 64 == 0).
constexpr int kGemmFwdW8 = 4;
constexpr int kGemmRhChoices[2] = {1, 2};

// One mirror pair of rings (north t, south nring-1-t); the equator of a grid
// with an odd number of rings is a pair with north == south (single = 1).
struct RingPair {
  int nphi;       // pixels on each ring of the pair (equal by symmetry)
  int M;          // FFT length: nphi when nphi is a power of two, else the
                  // Bluestein length pow2 >= 2*nphi - 1 (pow2 >= nphi - 1 for
                  // half pairs)
  int log2M;
  int single;     // 1 for the self-mirrored equator ring
  int64_t off_n;  // first pixel of the north ring
  int64_t off_s;  // first pixel of the south ring
  int chirp_off;  // Bluestein chirp w_j = exp(-i pi j^2 / nphi), nphi entries
  int bhat_off;   // Bluestein kernel spectrum FFT_M(conj w)/M, M entries
  float weight;   // ring quadrature weight
  int mwrite;     // the analysis writes the spectra for m < mwrite (what the
                  // forward GEMM reads of this pair; zero from min(mmax, nphi/2+1))
  int half;       // 1: each ring on its own as nphi/2 complex points
                  // z_j = f_2j + i f_2j+1; 2: z as its even and odd halves,
                  // two sequences of nphi/4 points (0: the pair packed)
  int nseq;       // points per transformed sequence (nphi >> half); the
                  // Bluestein tables are built for it when M != nseq
};

struct ShtPlan {
  int nring = 0;
  int npair = 0;      // number of ring pairs = (nring + 1) / 2
  int kpad = 0;       // npair rounded up to the Legendre K tile (32)
  int lmax = 0, mmax = 0;
  int64_t npix = 0;
  int max_M = 1;      // largest FFT length (power of two)
  int log2_max_M = 0;

  // Host copies
  std::vector<RingPair> pairs;
  std::vector<int64_t> p_off;   // [mmax + 1] float offset of m's block in the P table
  std::vector<int> nl;          // [mmax * 2] rows of each parity

  // Device tables
  RingPair* d_pairs = nullptr;
  float2* d_phase = nullptr;    // [npair][mmax] exp(-i m phi0)
  float2* d_chirp = nullptr;
  float2* d_bhat = nullptr;
  float2* d_twiddle = nullptr;  // [max_M] exp(-2 pi i x / max_M)
  // Scalar plans generate P_lm inside the GEMMs (no table): segment s of
  // order m covers l = m + 16 s .. m + 16 s + 15; the FP64-computed pair
  // (P_{l0}, P_{l0+1}) of every segment and ring pair starts an FP32
  // three-term recurrence over the other 14 degrees.
  float2* d_ck = nullptr;       // checkpoints [seg_off[m] + s][kpad]: (P_{l0}, P_{l0+1})
  float2* d_ab = nullptr;       // recurrence coefficients [seg_off[m] + s][16]: (a_l, b_l)
  int* d_seg_off = nullptr;     // [mmax + 1] first segment of each order
  float* d_ct = nullptr;        // [kpad] cos(theta) of the north ring of each pair (0 beyond)
  std::vector<int> seg_off;     // host copy
  std::vector<double> h_ct, h_phi0, h_w;  // per ring pair: cos(theta), phi0, weight (host, FP64)
  double* d_w64 = nullptr;                 // ring-pair weights (FP64 transforms)
  // FP64 transforms (hpx_sht_plan_enable_f64): the Legendre table in FP64 in
  // the layout of the FP32 table (p_off), Bluestein tables, twiddles and
  // phase factors in FP64.
  double* d_p64 = nullptr;
  RingPair* d_pairs64 = nullptr;  // ring pairs with offsets into the FP64 Bluestein tables
  double2* d_chirp64 = nullptr;
  double2* d_bhat64 = nullptr;
  double2* d_tw64 = nullptr;
  double2* d_phase64 = nullptr;
  int2* d_items64 = nullptr;      // forward FP64 GEMM items (m, parity * 65536 + row tile)
  int n_items64 = 0;
  int max_M64 = 1;                // largest FP64 FFT length (the FP64 pairs' own M)
  int log2_max_M64 = 0;
  int f64 = 0;
  // Vector plans hold tables instead:
  float* d_p = nullptr;         // Legendre table [m][parity][row][kpad]
  float* d_pt = nullptr;        // the same transposed per (m, parity): [m][parity][k][row]
  int64_t* d_p_off = nullptr;   // [mmax + 1]
  int* d_nl = nullptr;          // [mmax * 2]
  int2* d_fwd_items = nullptr;  // forward GEMM work items (m, row block)
  int* d_kmin = nullptr;        // [mmax] first ring pair whose spectrum reaches order m
  int* d_plist = nullptr;       // ring pairs grouped by FFT size class
  int class_off[4] = {0, 0, 0, 0};
  int class_n[4] = {0, 0, 0, 0};
  int n_fwd_items = 0;
  int64_t at_rows = 0;  // rows of the P table = rows of the coefficient workspace aT
  int64_t table_bytes = 0;
  // Vector plans (VectorSHT / VectoriSHT) hold two tables in the layout of P,
  // D0 (theta-derivative) at d_p and D1 (m P / sin theta) at d_p + vstride
  // (the same for d_pt), instead of P.
  int vector = 0;
  int csphase = 1;
  int vnorm = 0;     // HPX_NORM_* of the vector tables
  int vinverse = 0;  // vector tables scaled for the inverse transform
  int64_t vstride = 0;

  // Autotuned launch shapes (hpx_sht_autotune); untuned keys use defaults.
  mutable std::mutex tune_mu;
  std::map<int64_t, FftTune> fft_tunes;                   // by column count
  std::map<std::pair<int, int64_t>, GemmTune> gemm_tunes;  // by (precision, column count)

  FftTune fft_tune(int64_t ncol) const {
    std::lock_guard<std::mutex> g(tune_mu);
    auto it = fft_tunes.find(ncol);
    return it == fft_tunes.end() ? FftTune{} : it->second;
  }
  GemmTune gemm_tune(int precision, int64_t ncol) const {
    std::lock_guard<std::mutex> g(tune_mu);
    auto it = gemm_tunes.find({precision, ncol});
    if (it != gemm_tunes.end()) return it->second;
    return GemmTune{};
  }
};

// Rows of the P table per (m, parity) are padded to this many.
constexpr int kRowPad = 16;
// Degrees per recurrence segment of the scalar GEMMs (8 rows of each parity).
constexpr int kSegL = 16;
inline int nseg_of(int lmax, int m) { return m < lmax ? (lmax - m + kSegL - 1) / kSegL : 0; }
// Ring-pair (K) tile of the Legendre GEMMs.
constexpr int kKTile = 32;
// Largest FFT length (complex points) of the ring FFT kernels: 128 KB of
// LDS, the 16384-point Bluestein rings of nside 2048.
constexpr int kFftLds = 16384;
// FFT size classes: log2 M <= 10, 11, 12 (1, 2 or 4 wavefronts per
// transform) and 13, 14 (one work-group per transform).
constexpr int kFftClasses = 4;
int fft_class_of(int log2M);

inline int64_t col_pad(int64_t nbatch) { return ((2 * nbatch + 31) / 32) * 32; }

// Workspace: X[mmax][2][kpad][ncol] float (ring-pair spectra, S and D),
// then aT[at_rows][ncol] float: the coefficients in P-table row order
// (m, parity, row i with l = m + parity + 2 i), fields (re, im) along the row.
inline int64_t x_floats(const ShtPlan& p, int64_t nbatch) {
  return (int64_t)p.mmax * 2 * p.kpad * col_pad(nbatch);
}
// Rows after the last row of aT that the inverse GEMM may read (zero).
constexpr int kAtSlackRows = 64;
inline int64_t workspace_floats(const ShtPlan& p, int64_t nbatch) {
  return x_floats(p, nbatch) + (p.at_rows + kAtSlackRows) * col_pad(nbatch);
}
// Most maps one call may carry: the Legendre GEMMs address one order's
// spectra (2 kpad rows) and coefficient rows (at most lmax + 2 kRowPad) of
// col_pad(nbatch) floats through 32-bit buffer offsets.
inline int64_t max_maps(const ShtPlan& p) {
  const int64_t rows = std::max<int64_t>(2 * (int64_t)p.kpad, (int64_t)p.lmax + 2 * kRowPad);
  const int64_t max_ncol = 0x7fffffff / (rows * (int64_t)sizeof(float)) / 32 * 32;
  return max_ncol / 2;
}

// In-place radix-2 FFT on the host (FP64), exp(-2 pi i jk / n), n a power of two.
void host_fft(std::vector<std::complex<double>>& a);

// Launchers (return HPX_OK or an error code after setting the error text).
// layout 0: folded spectra workspace X[m][parity][kpad][ncol];
// layout 1: ring-major complex ftm[nbatch][nring][mmax] (no folding).
int launch_analysis(const ShtPlan& p, const float* f, float* X, int64_t nbatch,
                    int use_weight, float scale, int cmode, hipStream_t s, int layout = 0);
int launch_synthesis(const ShtPlan& p, const float* X, float* f, int64_t nbatch,
                     int use_weight, float scale, int cmode, hipStream_t s, int layout = 0);
// LDS bytes of a wave-level FFT work-group of size class c (< 3) with
// `waves` fields (their sequences and the pair's mmax phase factors).
size_t fft_lds_bytes(const ShtPlan& p, int c, int waves);
// One size class of the ring FFT stage with an explicit work-group shape.
int launch_fft_class(const ShtPlan& p, bool analysis, int cls, int waves, int tpf, const float* in,
                     float* out, int64_t nbatch, int use_weight, float scale, int cmode,
                     hipStream_t s, int layout);
// The GEMMs alone (X <-> aT) with an explicit column tiling, and the
// coefficient layout changes aT <-> a[b][l][m].
int launch_gemm_fwd(const ShtPlan& p, int precision, int ct, const float* X, float* aT,
                    int64_t nbatch, hipStream_t s);
int launch_gemm_inv(const ShtPlan& p, int precision, int ct, int rh, const float* aT, float* X,
                    int64_t nbatch, hipStream_t s);
int launch_coeff_to_user(const ShtPlan& p, const float* aT, float* alm, int64_t nbatch,
                         hipStream_t s);
int launch_user_to_coeff(const ShtPlan& p, const float* alm, float* aT, int64_t nbatch,
                         hipStream_t s);
// The Legendre launchers take the whole workspace (X, then aT).
int launch_legendre_fwd(const ShtPlan& p, int precision, float* ws, float* alm,
                        int64_t nbatch, hipStream_t s);
int launch_legendre_inv(const ShtPlan& p, int precision, const float* alm, float* ws,
                        int64_t nbatch, hipStream_t s);
int launch_legendre_table(ShtPlan& p, const double* theta_north, hipStream_t s);
// The two vector tables in FP64 (layout of the FP32 vector tables, D1 at
// D + vstride), for the FP64 vector transform.
int launch_vector_table64(const ShtPlan& p, double* D, const double* cos_theta, hipStream_t s);

}  // namespace hpxmi
