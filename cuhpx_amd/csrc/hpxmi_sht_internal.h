// Internal plan layout and kernel launchers of the SHT (not part of the C ABI).
#pragma once

#include "hpxmi_common.h"

#include <vector>

namespace hpxmi {

// One mirror pair of rings (north t, south nring-1-t); the equator of a grid
// with an odd number of rings is a pair with north == south (single = 1).
struct RingPair {
  int nphi;       // pixels on each ring of the pair (equal by symmetry)
  int M;          // FFT length: nphi when nphi is a power of two, else the
                  // Bluestein length pow2 >= 2*nphi - 1
  int log2M;
  int single;     // 1 for the self-mirrored equator ring
  int64_t off_n;  // first pixel of the north ring
  int64_t off_s;  // first pixel of the south ring
  int chirp_off;  // Bluestein chirp w_j = exp(-i pi j^2 / nphi), nphi entries
  int bhat_off;   // Bluestein kernel spectrum FFT_M(conj w)/M, M entries
  float weight;   // ring quadrature weight
  float pad_;
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
  float* d_p = nullptr;         // Legendre table [m][parity][row][kpad]
  int64_t* d_p_off = nullptr;   // [mmax + 1]
  int* d_nl = nullptr;          // [mmax * 2]
  int2* d_fwd_items = nullptr;  // forward GEMM work items (m, row block)
  int* d_plist = nullptr;       // ring pairs grouped by FFT size class
  int class_off[4] = {0, 0, 0, 0};
  int class_n[4] = {0, 0, 0, 0};
  int n_fwd_items = 0;
  int64_t table_bytes = 0;
};

// Rows of the P table per (m, parity) are padded to this many.
constexpr int kRowPad = 16;
// Ring-pair (K) tile of the Legendre GEMMs.
constexpr int kKTile = 32;
// Columns (2 * fields: re, im) per Legendre work-group.
constexpr int kColBlock = 128;
// Fields per FFT work-group.
constexpr int kFieldsPerFftBlock = 8;
// Largest FFT length (complex points) of the ring FFT kernels.
constexpr int kFftLds = 8192;
// FFT size classes: log2 M <= 10, 11, 12 (one wavefront per transform) and
// 13 (one work-group per transform).
constexpr int kFftClasses = 4;
int fft_class_of(int log2M);

inline int64_t col_pad(int64_t nbatch) { return ((2 * nbatch + 31) / 32) * 32; }

// Workspace: X[mmax][2][kpad][ncol_pad] float (ring-pair spectra, S and D).
inline int64_t workspace_floats(const ShtPlan& p, int64_t nbatch) {
  return (int64_t)p.mmax * 2 * p.kpad * col_pad(nbatch);
}

// Launchers (return HPX_OK or an error code after setting the error text).
// layout 0: folded spectra workspace X[m][parity][kpad][ncol];
// layout 1: ring-major complex ftm[nbatch][nring][mmax] (no folding).
int launch_analysis(const ShtPlan& p, const float* f, float* X, int64_t nbatch,
                    int use_weight, float scale, int cmode, hipStream_t s, int layout = 0);
int launch_synthesis(const ShtPlan& p, const float* X, float* f, int64_t nbatch,
                     int use_weight, float scale, int cmode, hipStream_t s, int layout = 0);
int launch_legendre_fwd(const ShtPlan& p, int precision, const float* X, float* alm,
                        int64_t nbatch, hipStream_t s);
int launch_legendre_inv(const ShtPlan& p, int precision, const float* alm, float* X,
                        int64_t nbatch, hipStream_t s);
int launch_legendre_table(ShtPlan& p, const double* theta_north, hipStream_t s);

}  // namespace hpxmi
