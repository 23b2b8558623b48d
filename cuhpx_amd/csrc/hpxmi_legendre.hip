// Associated-Legendre step of the SHT on MFMA (gfx950).
//
// Replaces the reference's dense einsums (cuhpx/hpx_sht.py:557 forward,
// :599 inverse, and the same contractions in the backward passes 573, 638),
// which run one FP32 GEMM per m over all l and all rings, including the zero
// triangle l < m, after a fresh pct*weights product on every call (546).
//
// MI355X design:
//  * Equatorial folding.  P_lm(-z) = (-1)^(l+m) P_lm(z), so with the ring
//    pair spectra S = F_t + F_t' and D = F_t - F_t' produced by the FFT stage,
//      a[l, m] = sum_{k < npair} P_lm(z_k) (l+m even ? S_k : D_k)[m]
//    and, inversely, Y_p[k] = sum_{l: l+m = p mod 2} P_lm(z_k) a[l, m].
//    This halves both the flops and the table, and only l >= m is touched.
//  * Table P[m][parity][i][k] (row i is l = m + parity + 2i) is generated on
//    the device in FP64 by the three-term recurrence with a power-of-two
//    exponent carried beside each value (no underflow to zero while the value
//    can still matter), rounded once to FP32.
//  * GEMM on MFMA: v_mfma_f32_16x16x4_f32 for the FP32 path, and
//    v_mfma_f32_16x16x32_bf16 for the BF16 paths.  The 3-term split product
//    hi*hi + hi*lo + lo*hi (x = hi + lo, both BF16, split from the FP32
//    operands; the lo*lo term is below FP32 rounding) accumulates in FP32.
//  * Forward work-group: one m, 64 rows of each parity, 64 or 128 columns
//    (re/im of 32 or 64 fields; 16 or 32 per wave).  The table rows are
//    staged once per k chunk through LDS and shared by the 4 waves; each wave
//    streams its columns of the spectra straight into MFMA B registers.  The
//    output aT[m][parity][i][col] leaves through LDS as whole row segments;
//    coeff_to_user_kernel turns it into the user layout a[b][l][m].
//  * Inverse work-group: one m, 64 (or 128) ring pairs, 64 or 128 columns;
//    A tiles come from the transposed table PT[m][parity][k][i], built once
//    at plan time, so they stage into LDS without a transposition.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

#include <algorithm>

namespace hpxmi {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kRowsPerParity = 64;  // forward: rows i per parity per work-group
constexpr int kKBlock = 64;         // inverse: ring pairs per work-group
// FP32 tile rows of 36 floats: row i starts at bank 36 i mod 64, sixteen
// distinct multiples of 4, so the fragment reads (row lane & 15, column
// 4 q + lane >> 4) hit all 64 banks once, and rows stay 16-byte aligned for
// the staging stores (33-float rows: 2-way conflicts, 4-byte stores).
constexpr int kLdsPadF = 4;

__device__ __forceinline__ void split(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// split() of two values at once, as packed BF16 pairs (value a in the low
// half): hi = RNE(a, b) by one v_cvt_pk_bf16_f32, its halves widened back by
// a shift and a mask, lo = RNE(a - hi_a, b - hi_b) by a second conversion:
// 3 VALU per value where the compiler makes 4 of split() (bit-identical).
__device__ __forceinline__ void split2(float a, float b, uint32_t& hi, uint32_t& lo) {
  const floatx2 v = {a, b};
  hi = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
  const floatx2 r = {a - __builtin_bit_cast(float, hi << 16), b - __builtin_bit_cast(float, hi & 0xffff0000u)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf16x2));
}

// ---------------------------------------------------------------------------
// Table generation (FP64, scaled recurrence)
// ---------------------------------------------------------------------------

// One thread per (m, k): P_mm by the sectoral product, then the three-term
// recurrence in l (cuhpx/sht_tools.py:243-252), in FP64, each value carried
// as v * 2^(600 e) (a value is zero while e < 0: |P| < 2^-600, far below
// FP32's range).  At the first degree l0 = m + 16 s of every segment s the
// pair (P_{l0}, P_{l0+1}) is stored, rounded once to FP32; the GEMMs
// regenerate the segment's other 14 degrees from it in FP32.
__global__ __launch_bounds__(kThreads) void checkpoint_kernel(float2* __restrict__ ck,
                                                              const int* __restrict__ seg_off,
                                                              const double* __restrict__ ct, int npair,
                                                              int kpad, int lmax) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= kpad) return;
  const int ns = seg_off[m + 1] - seg_off[m];
  float2* C = ck + (int64_t)seg_off[m] * kpad + k;
  if (k >= npair) {
    for (int t = 0; t < ns; ++t) C[(int64_t)t * kpad] = make_float2(0.f, 0.f);
    return;
  }
  if (ns == 0) return;
  const double x = ct[k];
  const double sn = sqrt((1.0 - x) * (1.0 + x));
  const double big = 0x1p600, tiny = 0x1p-600;
  double pmm = 0.28209479177387814347;  // 1 / sqrt(4 pi)
  int e = 0;
  for (int i = 1; i <= m; ++i) {
    pmm *= sqrt((2.0 * i + 1.0) / (2.0 * i)) * sn;
    if (pmm != 0.0 && fabs(pmm) < tiny) {
      pmm *= big;
      --e;
    }
  }
  // invariant at the top of the loop: p0 = P_l, p1 = P_{l+1} (both scaled by 2^(600 e))
  double p0 = pmm;
  double p1 = m + 1 < lmax ? sqrt(2.0 * m + 3.0) * x * pmm : 0.0;
  if (e < 0 && fabs(p1) > 1.0) {
    p0 *= tiny;
    p1 *= tiny;
    ++e;
  }
  const double dm = m;
  for (int l = m; l < lmax; ++l) {
    if (((l - m) & (kSegL - 1)) == 0)
      C[(int64_t)((l - m) / kSegL) * kpad] =
          e == 0 ? make_float2((float)p0, l + 1 < lmax ? (float)p1 : 0.f) : make_float2(0.f, 0.f);
    double v = 0.0;
    if (l + 2 < lmax) {
      const double dl = l + 2;
      const double a = sqrt((2 * dl - 1) / (dl - dm) * (2 * dl + 1) / (dl + dm));
      const double b =
          sqrt((dl + dm - 1) / (dl - dm) * (2 * dl + 1) / (2 * dl - 3) * (dl - dm - 1) / (dl + dm));
      v = x * a * p1 - b * p0;
    }
    p0 = p1;
    p1 = v;
    if (e < 0 && fabs(p1) > 1.0) {
      p0 *= tiny;
      p1 *= tiny;
      ++e;
    }
  }
}

// One segment of P_lm for one ring pair (cos theta = x): v[j] = P_{l0 + j},
// j < 16, from the stored pair v[0], v[1] by
// P_l = a_l x P_{l-1} - b_l P_{l-2} in FP32 (ab[j] = (a, b) of l0 + j; zero
// from lmax on, so the degrees past the band limit come out zero).
template <class AB>
__device__ __forceinline__ void segment_recurrence(float (&v)[kSegL], float2 c, float x, AB ab) {
  v[0] = c.x;
  v[1] = c.y;
#pragma unroll
  for (int j = 2; j < kSegL; ++j) {
    const float2 q = ab(j);
    v[j] = fmaf(q.x * x, v[j - 1], -(q.y * v[j - 2]));
  }
}

// P^mu_L(x) for L = mu, mu + 1, ... by the recurrence above, two consecutive
// degrees at a time (v0 = P_L, v1 = P_{L+1}), carried as v * 2^(600 e).
struct ScaledOrder {
  double v0, v1;
  int e, L, mu;

  __device__ void init(int order, double x, double s) {
    const double big = 0x1p600, tiny = 0x1p-600;
    mu = order;
    L = order;
    e = 0;
    double pmm = 0.28209479177387814347;  // 1 / sqrt(4 pi)
    for (int i = 1; i <= order; ++i) {
      pmm *= sqrt((2.0 * i + 1.0) / (2.0 * i)) * s;
      if (pmm != 0.0 && fabs(pmm) < tiny) {
        pmm *= big;
        --e;
      }
    }
    v0 = pmm;
    v1 = sqrt(2.0 * order + 3.0) * x * pmm;
    if (e < 0 && fabs(v1) > 1.0) {
      v0 *= tiny;
      v1 *= tiny;
      ++e;
    }
  }
  // P^mu_l for l in {L, L + 1}; zero below the order and while scaled down
  __device__ double at(int l) const {
    if (l < mu || e != 0) return 0.0;
    return l == L ? v0 : v1;
  }
  __device__ void advance(double x) {
    const double dl = L + 2, dm = mu;
    const double a = sqrt((2 * dl - 1) / (dl - dm) * (2 * dl + 1) / (dl + dm));
    const double b = sqrt((dl + dm - 1) / (dl - dm) * (2 * dl + 1) / (2 * dl - 3) * (dl - dm - 1) / (dl + dm));
    const double v = x * a * v1 - b * v0;
    v0 = v1;
    v1 = v;
    ++L;
    if (e < 0 && fabs(v1) > 1.0) {
      v0 *= 0x1p-600;
      v1 *= 0x1p-600;
      ++e;
    }
  }
};

// The two tables of the vector transform (sht_tools.py:267-304), one thread
// per (m, ring pair k), in the layout of the scalar table: D0 at D, D1 at
// D + vstride.
//   D0[l] = 1/2 (sqrt((l+m)(l-m+1)) P^{m-1}_l - sqrt((l-m)(l+m+1)) P^{m+1}_l)
//           (m = 0: -sqrt(l(l+1)) P^1_l), the theta-derivative of P^m_l;
//   D1[l] = 1/2 sqrt((2l+1)/(2l+3)) (sqrt((l-m+1)(l-m+2)) P^{m-1}_{l+1}
//           + sqrt((l+m+1)(l+m+2)) P^{m+1}_{l+1})  (m >= 1), = m P^m_l / sin;
// both negated for odd m when csphase (the Condon-Shortley sign).  Only the
// orders m - 1 and m + 1 are needed; each runs its own scaled recurrence.
// norm (sht_tools.py:194-221, applied to the P rows the two tables are built
// from): HPX_NORM_ORTHO leaves them; any other scales both by sqrt(4 pi),
// and HPX_NORM_SCHMIDT by (2 L + 1)^(-1/2) at the degree L the row reads
// (L = l for D0, l + 1 for D1); inverse tables take the reciprocals.
// T: float (the FP32 paths) or double (the FP64 vector transform).
template <typename T>
__global__ __launch_bounds__(kThreads) void vector_table_kernel(
    T* __restrict__ D, int64_t vstride, const int64_t* __restrict__ p_off, const int* __restrict__ nl,
    const double* __restrict__ ct, int npair, int kpad, int lmax, int csphase, int norm, int inverse) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= kpad) return;
  const int nl0 = nl[2 * m], nl1 = nl[2 * m + 1];
  const int pad0 = (nl0 + kRowPad - 1) / kRowPad * kRowPad;
  const int pad1 = (nl1 + kRowPad - 1) / kRowPad * kRowPad;
  T* T0[2] = {D + p_off[m], D + p_off[m] + (int64_t)pad0 * kpad};
  T* T1[2] = {T0[0] + vstride, T0[1] + vstride};
  for (int i = nl0; i < pad0; ++i) T0[0][(int64_t)i * kpad + k] = T1[0][(int64_t)i * kpad + k] = T(0);
  for (int i = nl1; i < pad1; ++i) T0[1][(int64_t)i * kpad + k] = T1[1][(int64_t)i * kpad + k] = T(0);
  if (k >= npair) {
    for (int i = 0; i < nl0; ++i) T0[0][(int64_t)i * kpad + k] = T1[0][(int64_t)i * kpad + k] = T(0);
    for (int i = 0; i < nl1; ++i) T0[1][(int64_t)i * kpad + k] = T1[1][(int64_t)i * kpad + k] = T(0);
    return;
  }
  if (m >= lmax) return;
  const double x = ct[k];
  const double s = sqrt((1.0 - x) * (1.0 + x));
  const T sign = (csphase && (m & 1)) ? T(-1) : T(1);
  const double dm = m;
  ScaledOrder lo, hi;  // orders m - 1 and m + 1
  hi.init(m + 1, x, s);
  if (m > 0) {
    lo.init(m - 1, x, s);
    lo.advance(x);  // degrees m, m + 1
  }
  for (int l = m; l < lmax; ++l) {
    const double dl = l;
    double d0, d1;
    if (m == 0) {
      d0 = -sqrt(dl * (dl + 1)) * hi.at(l);
      d1 = 0.0;
    } else {
      d0 = 0.5 * (sqrt((dl + dm) * (dl - dm + 1)) * lo.at(l) - sqrt((dl - dm) * (dl + dm + 1)) * hi.at(l));
      d1 = 0.5 * sqrt((2 * dl + 1) / (2 * dl + 3)) *
           (sqrt((dl - dm + 1) * (dl - dm + 2)) * lo.at(l + 1) + sqrt((dl + dm + 1) * (dl + dm + 2)) * hi.at(l + 1));
      lo.advance(x);
    }
    if (l > m) hi.advance(x);  // keep degrees l + 1, l + 2 for the next l
    if (norm != HPX_NORM_ORTHO) {
      const double c = 3.5449077018110320546;  // sqrt(4 pi)
      double f0 = c, f1 = c;
      if (norm == HPX_NORM_SCHMIDT) {
        f0 /= sqrt(2.0 * dl + 1.0);
        f1 /= sqrt(2.0 * dl + 3.0);
      }
      d0 = inverse ? d0 / f0 : d0 * f0;
      d1 = inverse ? d1 / f1 : d1 * f1;
    }
    const int par = (l - m) & 1;
    const int64_t off = (int64_t)((l - m) >> 1) * kpad + k;
    T0[par][off] = sign * (T)d0;
    T1[par][off] = sign * (T)d1;
  }
}

// ---------------------------------------------------------------------------
// Forward: aT[m][p][i][col] = sum_k P[m][p][i][k] X[m][p][k][col]
// ---------------------------------------------------------------------------

// Double-buffered table tile of the GEMMs: ROWS rows of a kKTile chunk, per
// parity, FP32 or split into BF16 hi / lo.
template <int PREC, int ROWS>
struct GemmLds;

template <int ROWS>
struct GemmLds<HPX_PREC_FP32, ROWS> {
  float a[2][ROWS][kKTile + kLdsPadF];
};

// BF16 tiles: unpadded 64-byte rows of 4 chunks (8 values, 16 bytes each);
// chunk c of row r sits in slot c ^ 2 ((r >> 3) & 1).  A 16-byte LDS read
// is served in groups of 16 lanes made of lane quads 0-3, 8-11, 20-23,
// 28-31 (and the like); with the A-fragment map (row lane & 15, chunk
// lane >> 4) this slot order puts each group on all 64 banks once.  Measured
// on the fragment read pattern alone: no bank conflicts and 1.7x the read
// rate of 80-byte padded rows (4 conflict cycles per read).  In the GEMMs it
// halves the conflict cycles (the staging stores keep theirs) and leaves the
// time unchanged, the loads from HBM being the limit; it also cuts a
// work-group's tile buffers from 40 to 32 KB.
template <int ROWS>
struct GemmLds<HPX_PREC_BF16X3, ROWS> {
  __bf16 hi[2][ROWS][kKTile];
  __bf16 lo[2][ROWS][kKTile];
};

template <int ROWS>
struct GemmLds<HPX_PREC_BF16, ROWS> {
  __bf16 hi[2][ROWS][kKTile];
};

// Element offset of value j of row r in a BF16 tile row (XOR on chunk bit 1).
__device__ __forceinline__ int bf16_slot(int r, int j) { return j ^ (((r >> 3) & 1) << 4); }

template <int PREC>
using FwdLds = GemmLds<PREC, kRowsPerParity>;

// Split a value into BF16 hi and lo (x = hi + lo to ~16 bits more).
template <int PREC>
__device__ __forceinline__ void split_to(const float (&v)[8], bf16x8& hi, bf16x8& lo) {
  if constexpr (PREC == HPX_PREC_BF16X3) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) split2(v[2 * q], v[2 * q + 1], h[q], l[q]);
    hi = __builtin_bit_cast(bf16x8, h);
    lo = __builtin_bit_cast(bf16x8, l);
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) hi[q] = (__bf16)v[q];
  }
}

// One k (or i) chunk of MFMA products for 2 parities x 4 row tiles x CT
// column tiles: A fragments from LDS (row-major tile, row = lane & 15),
// B fragments from registers.  Straight-line: row tiles beyond the valid
// rows multiply zeros.
template <int PREC, int CT, class LdsT>
__device__ __forceinline__ void mfma_chunk(const LdsT& L, const float (&bv)[2][CT][8], floatx4 (&acc)[2][4][CT],
                                           int lane, int row_off = 0) {
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const int ar = row_off + rt * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const float av = L.a[p][ar][4 * q + (lane >> 4)];
#pragma unroll
          for (int ct = 0; ct < CT; ++ct)
            acc[p][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[p][ct][q], acc[p][rt][ct], 0, 0, 0);
        }
      }
    } else {
      bf16x8 bhi[CT], blo[CT];
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) split_to<PREC>(bv[p][ct], bhi[ct], blo[ct]);
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const int ar = row_off + rt * 16 + (lane & 15);
        const int aj = bf16_slot(ar, 8 * (lane >> 4));
        const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(&L.hi[p][ar][aj]);
        bf16x8 alo;
        if constexpr (PREC == HPX_PREC_BF16X3)
          alo = *reinterpret_cast<const bf16x8*>(&L.lo[p][ar][aj]);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
          acc[p][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bhi[ct], acc[p][rt][ct], 0, 0, 0);
          if constexpr (PREC == HPX_PREC_BF16X3) {
            acc[p][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, blo[ct], acc[p][rt][ct], 0, 0, 0);
            acc[p][rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bhi[ct], acc[p][rt][ct], 0, 0, 0);
          }
        }
      }
    }
  }
}

// k offset (within a chunk) of B value q of a lane: the MFMA B fragment
// holds k = 4q + lane/16 (FP32, 16x16x4) or k = 8 (lane/16) + q (BF16, 16x16x32).
template <int PREC>
__device__ __forceinline__ int frag_k(int q, int lane) {
  return PREC == HPX_PREC_FP32 ? 4 * q + (lane >> 4) : 8 * (lane >> 4) + q;
}
// The same split into a per-lane part and a wave-uniform part.
template <int PREC>
__device__ __forceinline__ int frag_k_lane(int lane) {
  return PREC == HPX_PREC_FP32 ? (lane >> 4) : 8 * (lane >> 4);
}
template <int PREC>
constexpr int frag_k_uniform(int q) {
  return PREC == HPX_PREC_FP32 ? 4 * q : q;
}

// Raw buffer access (gfx9 descriptor: 32-bit dwords, bounds-checked against
// nbytes, out-of-range loads return 0).  B fragments are strided loads whose
// offsets are wave-uniform except for one lane term: the uniform part goes in
// the scalar offset, so a chunk's loads need one vector register for address.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const float* base, int64_t nbytes) {
  const int n = nbytes > 0x7fffffff ? 0x7fffffff : (int)nbytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), (short)0, n, 0x00020000);
}
__device__ __forceinline__ float buffer_load(__amdgpu_buffer_rsrc_t r, int voff_bytes, int soff_bytes) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff_bytes, soff_bytes, 0));
}

// Write the two 64-row x (64 CT)-column accumulator tiles of a work-group
// (one per parity) through LDS, so that each store instruction writes whole
// row segments, 16 bytes per lane.  T: LDS scratch of at least
// 64 * (64 CT + 4) floats, 16-byte aligned (the table buffers, free after
// the main loop); row(p, r): destination of local row r of parity p, or
// nullptr; only the first ncol_left columns of a row exist.  All threads of
// the work-group call it.
// NT: non-temporal stores (the inverse GEMM's 269 MB of spectra at nside 256,
// 64 fields, streamed out; measured neutral to 2% faster).
template <int CT, int RH = 1, bool NT = false, class RowFn>
__device__ __forceinline__ void store_tiles_lds(const floatx4 (&acc)[2][4][CT], float* T, int wave, int lane,
                                                int64_t ncol_left, RowFn row) {
  // RH row halves of 64 rows; wave = 4 * row half + column wave
  constexpr int kW = 64 * CT;
  constexpr int kS = kW + 4;
  const int rh = wave >> 2;
  const int cw = wave & 3;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          T[(rh * 64 + rt * 16 + 4 * (lane >> 4) + r) * kS + cw * 16 * CT + ct * 16 + (lane & 15)] =
              acc[p][rt][ct][r];
    __syncthreads();
    for (int v = threadIdx.x; v < 64 * RH * (kW / 4); v += kThreads * RH) {
      const int r = v / (kW / 4);
      const int c4 = v - r * (kW / 4);
      float* dst = row(p, r);
      if (dst != nullptr && 4 * c4 < ncol_left)
      {
        const float4 val = *reinterpret_cast<const float4*>(&T[r * kS + 4 * c4]);
        if constexpr (NT) {
          float* q = dst + 4 * c4;
          __builtin_nontemporal_store(val.x, q);
          __builtin_nontemporal_store(val.y, q + 1);
          __builtin_nontemporal_store(val.z, q + 2);
          __builtin_nontemporal_store(val.w, q + 3);
        } else {
          *reinterpret_cast<float4*>(dst + 4 * c4) = val;
        }
      }
    }
  }
}

// Occupancy target of the GEMM kernels (waves per SIMD) that their register
// use allows without spilling.
constexpr int gemm_min_waves(int prec, int ct) { return ct == 1 ? 4 : (prec == HPX_PREC_FP32 ? 3 : 2); }

// Lane signs and column swaps of the vector transform's GEMMs.  Columns are
// grouped by field in fours, q = column & 3 = (component, re / im) with
// component 0 = spheroidal / theta, 1 = toroidal / phi.  The reduction runs
// over two halves: the D0 table against spectra (coefficients) of the
// opposite parity block with column q and sign s0(q) = (+, +, -, -), then
// the D1 table against the same parity block with column q ^ 3 and sign
// s1(q) = (+, -, +, -) in the forward transform, (-, +, -, +) in the inverse
// (hpx_sht.py:351-372 and 413-432 written out per real column).
__device__ __forceinline__ float vec_sign0(int lane) { return (lane & 2) ? -1.f : 1.f; }
__device__ __forceinline__ float vec_sign1(int lane, bool forward) {
  return ((lane & 1) != 0) == forward ? -1.f : 1.f;
}

// What the scalar GEMMs generate P_lm from (the plan's segment checkpoints,
// recurrence coefficients and ring cosines).
struct OtfArgs {
  const float2* __restrict__ ck;
  const float2* __restrict__ ab;
  const int* __restrict__ seg_off;
  const float* __restrict__ ct;
};

// Scalar transform (VEC = false): the table tile of each k chunk is
// generated in the work-group: thread (segment s, ring pair k) runs the FP32
// recurrence of its 16 degrees (8 rows of each parity) from the segment's
// checkpoint and writes them into the tile's column k.
// VEC: the vector transform, tiles read from the stored tables; the
// reduction is 2 kpad long (D0 half, then D1 half at P + vstride), see
// vec_sign0 / vec_sign1.
template <int PREC, int CT, bool VEC = false>
__global__ __launch_bounds__(kThreads, gemm_min_waves(PREC, CT)) void legendre_fwd_kernel(
    const float* __restrict__ X, float* __restrict__ aT, const float* __restrict__ P, OtfArgs g,
    const int64_t* __restrict__ p_off, const int* __restrict__ nl, const int* __restrict__ kmin,
    const int2* __restrict__ items, int nitems, int ncb, int kpad, int64_t ncol, int64_t vstride) {
  // table tiles (double buffer) during the main loop, the output tile after it
  constexpr size_t kTileBytes = sizeof(float) * 64 * (64 * CT + 4);
  constexpr size_t kLdsBytes = 2 * sizeof(FwdLds<PREC>) > kTileBytes ? 2 * sizeof(FwdLds<PREC>) : kTileBytes;
  __shared__ __align__(16) unsigned char smem[kLdsBytes];
  FwdLds<PREC>* lds = reinterpret_cast<FwdLds<PREC>*>(smem);
  // work index = (item, column block), column block fastest: the column
  // blocks of one table tile run side by side on one XCD and share the tile
  const int idx = xcd_local_index(blockIdx.x, gridDim.x);
  if (idx >= nitems * ncb) return;
  const int item = idx / ncb;
  const int cb = idx - item * ncb;
  const int m = items[item].x;
  const int jb = items[item].y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // a wave whose columns lie beyond the batch loads and multiplies the block's
  // first columns instead and stores nothing, so that every wave issues the
  // same loads
  const int col0 = cb * (64 * CT) + wave * (16 * CT) < ncol ? cb * (64 * CT) + wave * (16 * CT) : 0;
  const int nlp[2] = {nl[2 * m], nl[2 * m + 1]};
  const int padp0 = (nlp[0] + kRowPad - 1) / kRowPad * kRowPad;
  const int padp1 = (nlp[1] + kRowPad - 1) / kRowPad * kRowPad;
  const int row0 = jb * kRowsPerParity;
  const float* Pm = P + p_off[m];
  const int64_t pstride = (int64_t)kpad * ncol;
  const int nc = (int)ncol;

  floatx4 acc[2][4][CT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[p][rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};

  // staging assignment: 2 parities x 64 rows x 2 halves of 16 k
  const int sp = threadIdx.x >> 7;
  const int sr = (threadIdx.x >> 1) & 63;
  const int sh = (threadIdx.x & 1) * 16;
  const int shs = bf16_slot(sr, sh);  // the same 16 values in a BF16 tile row
  const int srow = row0 + sr;
  const bool srow_ok = srow < (sp ? padp1 : padp0);
  const float* Psrc = Pm + (sp ? (int64_t)padp0 * kpad : 0) + (int64_t)srow * kpad + sh;
  // generation (scalar): 8 segments (this work-group's 128 degrees) x 32 ring
  // pairs; the segments' recurrence coefficients stay in LDS
  __shared__ float2 abl[VEC ? 1 : 8][kSegL];
  const int gs = threadIdx.x >> 5;
  const int gk = threadIdx.x & 31;
  const int sbase = VEC ? 0 : g.seg_off[m];
  const int nsg = VEC ? 0 : g.seg_off[m + 1] - sbase;
  // this order's checkpoints through a buffer descriptor: a segment past the
  // order's last lies beyond it and reads zero, with no branch
  const __amdgpu_buffer_rsrc_t ckr =
      buffer_rsrc(reinterpret_cast<const float*>(VEC ? nullptr : g.ck + (int64_t)sbase * kpad),
                  VEC ? 0 : (int64_t)nsg * kpad * 8);
  const int ckoff = ((jb * 8 + gs) * kpad + gk) * 8;
  if constexpr (!VEC) {
    if (threadIdx.x < 8 * kSegL) {
      const int ts = threadIdx.x / kSegL, tj = threadIdx.x % kSegL;
      abl[ts][tj] = jb * 8 + ts < nsg ? g.ab[(int64_t)(sbase + jb * 8 + ts) * kSegL + tj] : make_float2(0.f, 0.f);
    }
    __syncthreads();
  }
  // B fragments from X[m] (both parities): the rows beyond the last ring
  // pair are zero, so no k bound check
  const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + (int64_t)m * 2 * pstride, 2 * pstride * 4);
  const int bvoff = (frag_k_lane<PREC>(lane) * nc + col0 + (lane & 15)) * 4;
  const int bvoff3 = (frag_k_lane<PREC>(lane) * nc + col0 + ((lane & 15) ^ 3)) * 4;
  const float sgn0 = vec_sign0(lane), sgn1 = vec_sign1(lane, true);

  auto load_a = [&](int kc, float* v) {
    if constexpr (!VEC) {  // the checkpoint and the ring cosine; store_a generates
      v[0] = buffer_load(ckr, ckoff, kc * 8);
      v[1] = buffer_load(ckr, ckoff + 4, kc * 8);
      v[2] = g.ct[kc + gk];
      return;
    }
    // a row past the parity's count reads row 0 (valid memory) and is zeroed
    // in store_a: the load itself is unconditional
    const float* src = (srow_ok ? Psrc : Pm + sh) + kc;
    if (VEC && kc >= kpad) src += vstride - kpad;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 t = *reinterpret_cast<const float4*>(src + 4 * q);
      v[4 * q] = t.x;
      v[4 * q + 1] = t.y;
      v[4 * q + 2] = t.z;
      v[4 * q + 3] = t.w;
    }
  };
  auto store_a = [&](FwdLds<PREC>& L, const float* va) {
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = (VEC && !srow_ok) ? 0.f : va[q];
    if constexpr (!VEC) {
      float w[kSegL];
      segment_recurrence(w, make_float2(v[0], v[1]), v[2], [&](int j) { return abl[gs][j]; });
#pragma unroll
      for (int j = 0; j < kSegL; ++j) {
        const int p = j & 1, r = 8 * gs + (j >> 1);
        if constexpr (PREC == HPX_PREC_FP32) {
          L.a[p][r][gk] = w[j];
        } else if constexpr (PREC == HPX_PREC_BF16X3) {
          // pairs (j, j + 2): rows r and r + 1 of one parity
          if ((j & 2) == 0) {
            uint32_t h, l;
            split2(w[j], w[j + 2], h, l);
            L.hi[p][r][bf16_slot(r, gk)] = __builtin_bit_cast(bf16x2, h)[0];
            L.lo[p][r][bf16_slot(r, gk)] = __builtin_bit_cast(bf16x2, l)[0];
            L.hi[p][r + 1][bf16_slot(r + 1, gk)] = __builtin_bit_cast(bf16x2, h)[1];
            L.lo[p][r + 1][bf16_slot(r + 1, gk)] = __builtin_bit_cast(bf16x2, l)[1];
          }
        } else {
          L.hi[p][r][bf16_slot(r, gk)] = (__bf16)w[j];
        }
      }
      return;
    }
    if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.a[sp][sr][sh + q] = v[q];
    } else if constexpr (PREC == HPX_PREC_BF16X3) {
#pragma unroll
      for (int q = 0; q < 16; ++q) split(v[q], L.hi[sp][sr][shs + q], L.lo[sp][sr][shs + q]);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.hi[sp][sr][shs + q] = (__bf16)v[q];
    }
  };
  auto load_b = [&](int kc, float (&bv)[2][CT][8]) {
    const bool half1 = VEC && kc >= kpad;
    const int kk = half1 ? kc - kpad : kc;
    const int voff = half1 ? bvoff3 : bvoff;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int pb = VEC && !half1 ? 1 - p : p;  // spectra parity block
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          bv[p][ct][q] = buffer_load(xr, voff,
                                     (int)((pb * pstride + (int64_t)(kk + frag_k_uniform<PREC>(q)) * nc + ct * 16) * 4));
    }
  };
  // the vector transform's lane signs, applied just before the products
  // (writing a register that a load in flight targets would wait for it)
  auto prep_b = [&](int kc, float (&bv)[2][CT][8]) {
    if constexpr (VEC) {
      const float sg = kc >= kpad ? sgn1 : sgn0;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q) bv[p][ct][q] *= sg;
    }
  };

  // k chunks wholly below kmin[m] hold zero spectra and are skipped (in
  // both halves of a vector reduction); the others run in the same order,
  // so every sum is unchanged.  Chunk t of the loop starts at ring pair kof(t).
  const int c0 = min(kmin[m], kpad) / kKTile;
  const int half = kpad / kKTile - c0;
  const int nchunk = (VEC ? 2 : 1) * half;
  auto kof = [&](int t) { return (VEC && t >= half ? kpad + (t - half) * kKTile : t * kKTile) + c0 * kKTile; };

  // software pipeline over k chunks with two register sets for B and two
  // LDS buffers for the table tile (one barrier per chunk, no copies).  Every
  // half-iteration issues its loads unconditionally (the last one reloads
  // the last chunk, and its tile is never read): a load issued on only some
  // paths makes the compiler wait for every load in flight before the
  // products of the current chunk, which serialises the prefetch.
  float pa[16];
  float b0[2][CT][8], b1[2][CT][8];
  if (nchunk > 0) {
    load_a(kof(0), pa);
    load_b(kof(0), b0);
    store_a(lds[0], pa);
    __syncthreads();
    for (int c = 0;; c += 2) {
      const int c1 = min(c + 1, nchunk - 1);
      load_a(kof(c1), pa);
      load_b(kof(c1), b1);
      prep_b(kof(c), b0);
      mfma_chunk<PREC, CT>(lds[0], b0, acc, lane);
      store_a(lds[1], pa);
      __syncthreads();
      if (c + 1 >= nchunk) break;
      const int c2 = min(c + 2, nchunk - 1);
      load_a(kof(c2), pa);
      load_b(kof(c2), b0);
      prep_b(kof(c + 1), b1);
      mfma_chunk<PREC, CT>(lds[1], b1, acc, lane);
      store_a(lds[0], pa);
      __syncthreads();
      if (c + 2 >= nchunk) break;
    }
  }

  // ---- this m's rows of aT (rows row0.. of each parity, this work-group's columns)
  float* Ab = aT + (p_off[m] / kpad) * ncol + cb * (64 * CT);
  store_tiles_lds<CT>(acc, reinterpret_cast<float*>(smem), wave, lane, ncol - cb * (64 * CT),
                      [&](int p, int r) -> float* {
                        const int i = row0 + r;
                        return i < nlp[p] ? Ab + (int64_t)((p ? padp0 : 0) + i) * ncol : nullptr;
                      });
}

// Forward GEMM with 8 wavefronts (scalar transform): the work-group covers
// 128 columns (16 per wave) of 64 rows per parity, so one generated table
// tile serves all of them, and a k chunk is 64 ring pairs (two 32-pair
// halves, each a GemmLds tile), so every one of the 512 threads generates
// one 16-degree segment for one ring pair per chunk: per ring pair and
// column half the generation work and half the barriers of the 4-wave
// kernel, at the same occupancy (2 work-groups of 64 KB per CU, 4 waves per
// SIMD).  The spectra of a chunk's second half are loaded during the first
// half's products, the next chunk's first half during the second's.  Needs
// kpad % 64 == 0 (chunks start at multiples of 64, kmin rounded down).
constexpr int kW8Threads = 512;
template <int PREC>
__global__ __launch_bounds__(kW8Threads, 4) void legendre_fwd_w8_kernel(
    const float* __restrict__ X, float* __restrict__ aT, OtfArgs g, const int64_t* __restrict__ p_off,
    const int* __restrict__ nl, const int* __restrict__ kmin, const int2* __restrict__ items, int nitems,
    int ncb, int kpad, int64_t ncol) {
  using Half = GemmLds<PREC, kRowsPerParity>;  // one 32-ring-pair half of a chunk
  constexpr int kOutW = 128, kOutS = kOutW + 4;
  constexpr size_t kTileBytes = sizeof(float) * 64 * kOutS;
  constexpr size_t kLdsBytes = 4 * sizeof(Half) > kTileBytes ? 4 * sizeof(Half) : kTileBytes;
  __shared__ __align__(16) unsigned char smem[kLdsBytes];
  Half* lds = reinterpret_cast<Half*>(smem);  // lds[2 * buffer + half]
  __shared__ float2 abl[8][kSegL];
  const int idx = xcd_local_index(blockIdx.x, gridDim.x);
  if (idx >= nitems * ncb) return;
  const int item = idx / ncb;
  const int cb = idx - item * ncb;
  const int m = items[item].x;
  const int jb = items[item].y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  // a wave whose columns lie beyond the batch works on the first columns and
  // stores nothing, so that every wave issues the same loads
  const int col0 = cb * kOutW + wave * 16 < ncol ? cb * kOutW + wave * 16 : 0;
  const int nlp[2] = {nl[2 * m], nl[2 * m + 1]};
  const int padp0 = (nlp[0] + kRowPad - 1) / kRowPad * kRowPad;
  const int row0 = jb * kRowsPerParity;
  const int64_t pstride = (int64_t)kpad * ncol;
  const int nc = (int)ncol;

  floatx4 acc[2][4][1];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) acc[p][rt][0] = floatx4{0.f, 0.f, 0.f, 0.f};

  // generation: thread (segment gs of this work-group's 8, ring pair gk of
  // the chunk's 64; gh = gk / 32 the half it writes)
  const int gs = threadIdx.x >> 6;
  const int gk = threadIdx.x & 63;
  const int gh = gk >> 5, gkk = gk & 31;
  const int sbase = g.seg_off[m];
  const int nsg = g.seg_off[m + 1] - sbase;
  if (threadIdx.x < 8 * kSegL) {
    const int ts = threadIdx.x / kSegL, tj = threadIdx.x % kSegL;
    abl[ts][tj] = jb * 8 + ts < nsg ? g.ab[(int64_t)(sbase + jb * 8 + ts) * kSegL + tj] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const __amdgpu_buffer_rsrc_t ckr =
      buffer_rsrc(reinterpret_cast<const float*>(g.ck + (int64_t)sbase * kpad), (int64_t)nsg * kpad * 8);
  const int ckoff = ((jb * 8 + gs) * kpad + gk) * 8;
  const __amdgpu_buffer_rsrc_t xr = buffer_rsrc(X + (int64_t)m * 2 * pstride, 2 * pstride * 4);
  const int bvoff = (frag_k_lane<PREC>(lane) * nc + col0 + (lane & 15)) * 4;

  auto load_a = [&](int kc, float* v) {
    v[0] = buffer_load(ckr, ckoff, kc * 8);
    v[1] = buffer_load(ckr, ckoff + 4, kc * 8);
    v[2] = g.ct[kc + gk];
  };
  auto store_a = [&](int buf, const float* v) {
    Half& L = lds[2 * buf + gh];
    float w[kSegL];
    segment_recurrence(w, make_float2(v[0], v[1]), v[2], [&](int j) { return abl[gs][j]; });
#pragma unroll
    for (int j = 0; j < kSegL; ++j) {
      const int p = j & 1, r = 8 * gs + (j >> 1);
      if constexpr (PREC == HPX_PREC_FP32) {
        L.a[p][r][gkk] = w[j];
      } else if constexpr (PREC == HPX_PREC_BF16X3) {
        if ((j & 2) == 0) {
          uint32_t h, l;
          split2(w[j], w[j + 2], h, l);
          L.hi[p][r][bf16_slot(r, gkk)] = __builtin_bit_cast(bf16x2, h)[0];
          L.lo[p][r][bf16_slot(r, gkk)] = __builtin_bit_cast(bf16x2, l)[0];
          L.hi[p][r + 1][bf16_slot(r + 1, gkk)] = __builtin_bit_cast(bf16x2, h)[1];
          L.lo[p][r + 1][bf16_slot(r + 1, gkk)] = __builtin_bit_cast(bf16x2, l)[1];
        }
      } else {
        L.hi[p][r][bf16_slot(r, gkk)] = (__bf16)w[j];
      }
    }
  };
  auto load_b = [&](int kc, float (&bv)[2][1][8]) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int q = 0; q < 8; ++q)
        bv[p][0][q] = buffer_load(xr, bvoff, (int)((p * pstride + (int64_t)(kc + frag_k_uniform<PREC>(q)) * nc) * 4));
  };

  // chunks of 64 ring pairs from kmin[m] rounded down to 64.  The 32-pair
  // halves below kmin[m] rounded down to 32 (what the ring analysis leaves
  // unwritten for this m) are loaded but not multiplied, so the products
  // run over the same 32-pair chunks in the same order as in the 4-wave
  // kernel (bit-identical results).
  const int k32 = min(kmin[m], kpad) / kKTile * kKTile;
  const int c0 = k32 / 64;
  const int nchunk = kpad / 64 - c0;
  const bool skip_first_half = c0 * 64 < k32;
  auto kof = [&](int t) { return (c0 + t) * 64; };
  if (nchunk > 0) {
    float pa[3];
    float b0[2][1][8], b1[2][1][8];
    load_a(kof(0), pa);
    load_b(kof(0), b0);
    store_a(0, pa);
    __syncthreads();
    int cur = 0;
    for (int c = 0;; ++c) {
      const int cn = min(c + 1, nchunk - 1);
      load_a(kof(cn), pa);
      load_b(kof(c) + 32, b1);
      if (c > 0 || !skip_first_half) mfma_chunk<PREC, 1>(lds[2 * cur], b0, acc, lane);
      load_b(kof(cn), b0);
      mfma_chunk<PREC, 1>(lds[2 * cur + 1], b1, acc, lane);
      store_a(cur ^ 1, pa);
      __syncthreads();
      if (c + 1 >= nchunk) break;
      cur ^= 1;
    }
  }

  // output: both parities' 64 x 128 tiles through LDS, whole row segments
  float* T = reinterpret_cast<float*>(smem);
  float* Ab = aT + (p_off[m] / kpad) * ncol + cb * kOutW;
  const int64_t ncol_left = ncol - cb * kOutW;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        T[(rt * 16 + 4 * (lane >> 4) + r) * kOutS + wave * 16 + (lane & 15)] = acc[p][rt][0][r];
    __syncthreads();
    for (int v = threadIdx.x; v < 64 * (kOutW / 4); v += kW8Threads) {
      const int r = v / (kOutW / 4);
      const int c4 = v - r * (kOutW / 4);
      const int i = row0 + r;
      if (i < nlp[p] && 4 * c4 < ncol_left)
        *reinterpret_cast<float4*>(Ab + (int64_t)((p ? padp0 : 0) + i) * ncol + 4 * c4) =
            *reinterpret_cast<const float4*>(&T[r * kOutS + 4 * c4]);
    }
  }
}

// ---------------------------------------------------------------------------
// Inverse: Y[m][p][k][col] = sum_i P[m][p][i][k] aT[m][p][i][col]
// ---------------------------------------------------------------------------

// RH: row halves of 64 ring pairs per work-group (4 * RH waves: wave = 4 *
// row half + column wave).  With RH = 2 a work-group covers 128 ring pairs,
// so each coefficient chunk serves both halves and the per-work-group set-up
// and output staging are spread over twice the work.
// VEC: the vector transform; the reduction runs over the rows of both
// parities (D0T half against the opposite parity's coefficients, then the
// D1T half at PT + vstride against the same parity's), see vec_sign0.
// Scalar transform (VEC = false): the table tile of each i chunk is generated
// in the work-group: thread (ring pair k, segment g of the chunk's four) runs
// the FP32 recurrence of the segment's 16 degrees, which are 8 consecutive
// rows i of each parity, so they land in the tile as one 16-byte row piece
// per parity.
template <int PREC, int CT, int RH, bool VEC = false>
__global__ __launch_bounds__(kThreads * RH, CT == 1 ? 4 : 2) void legendre_inv_kernel(
    const float* __restrict__ aT, float* __restrict__ X, const float* __restrict__ PT, OtfArgs g,
    const int64_t* __restrict__ p_off, const int* __restrict__ nl, const int* __restrict__ kmin, int nkb,
    int nitems, int ncb, int kpad, int64_t ncol, int64_t at_rows, int64_t vstride) {
  // same tiling as the forward kernel with the roles of i and k exchanged:
  // rows = ring pairs, reduction over the rows i of aT, A tiles from the
  // transposed table PT[m][p][k][i]
  // table tiles (double buffer) during the main loop, the output tile after it
  constexpr int kRows = 64 * RH;
  using Lds = GemmLds<PREC, kRows>;
  constexpr size_t kTileBytes = sizeof(float) * kRows * (64 * CT + 4);
  constexpr size_t kLdsBytes = 2 * sizeof(Lds) > kTileBytes ? 2 * sizeof(Lds) : kTileBytes;
  __shared__ __align__(16) unsigned char smem[kLdsBytes];
  Lds* lds = reinterpret_cast<Lds*>(smem);
  // m outer, ring-pair block inner: the nkb work-groups of one m read the same
  // coefficient rows aT[m] and land on one XCD (contiguous item ranges).  The
  // reduction length of m is ~(mmax - m) / 2, so m runs in the order 0,
  // mmax-1, 1, mmax-2, ...: every contiguous eighth of the items (one XCD's
  // share) then carries the same total work.
  const int idx = xcd_local_index(blockIdx.x, gridDim.x);
  if (idx >= nitems * ncb) return;
  const int item = idx / ncb;
  const int cb = idx - item * ncb;
  const int mi = item / nkb;
  const int kb = item - mi * nkb;
  const int m = (mi & 1) ? nitems / nkb - 1 - (mi >> 1) : (mi >> 1);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int rh = wave >> 2;
  // a wave whose columns lie beyond the batch loads and multiplies the block's
  // first columns instead and stores nothing (every wave issues the same loads)
  const int col0 =
      cb * (64 * CT) + (wave & 3) * (16 * CT) < ncol ? cb * (64 * CT) + (wave & 3) * (16 * CT) : 0;
  const int nlp[2] = {nl[2 * m], nl[2 * m + 1]};
  const int padp[2] = {(nlp[0] + kRowPad - 1) / kRowPad * kRowPad, (nlp[1] + kRowPad - 1) / kRowPad * kRowPad};
  const int nmax = nlp[0] > nlp[1] ? nlp[0] : nlp[1];
  const int k0 = kb * kRows;
  // ring pairs below kmin[m] are read by no synthesis at this m (their
  // spectra stop below m): a block wholly below it has nothing to compute
  if (k0 + kRows <= kmin[m]) return;
  const int nc = (int)ncol;

  floatx4 acc[2][4][CT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[p][rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};

  // staging: 2 parities x kRows ring pairs x 2 halves of 16 i
  const int sp = threadIdx.x / (2 * kRows);
  const int sr = (threadIdx.x >> 1) & (kRows - 1);
  const int sh = (threadIdx.x & 1) * 16;
  const int shs = bf16_slot(sr, sh);  // the same 16 values in a BF16 tile row
  // first half: parity sp (scalar) or 1 - sp (vector, D0T); second half
  // (vector only): parity sp of D1T
  const int sp0 = VEC ? 1 - sp : sp;
  const int spad0 = padp[sp0], spad1 = padp[sp];
  const float* Pbase0 = PT + p_off[m] + (sp0 ? (int64_t)padp[0] * kpad : 0) + (int64_t)sr * spad0 + sh;
  const float* Pbase1 = PT + vstride + p_off[m] + (sp ? (int64_t)padp[0] * kpad : 0) + (int64_t)sr * spad1 + sh;
  const int64_t row0m = p_off[m] / kpad;
  const __amdgpu_buffer_rsrc_t ar = buffer_rsrc(aT + row0m * ncol, (at_rows + kAtSlackRows - row0m) * ncol * 4);
  const int bvoff = (frag_k_lane<PREC>(lane) * nc + col0 + (lane & 15)) * 4;
  const int bvoff3 = (frag_k_lane<PREC>(lane) * nc + col0 + ((lane & 15) ^ 3)) * 4;
  const float sgn0 = vec_sign0(lane), sgn1 = vec_sign1(lane, false);
  const int kp = (nmax + kKTile - 1) / kKTile * kKTile;  // length of one half
  // generation (scalar): ring pair gk of the block, segment gg of the chunk
  // (wave-uniform, so the recurrence coefficients are scalar loads)
  const int gk = threadIdx.x & (kRows - 1);
  const int gg = __builtin_amdgcn_readfirstlane(threadIdx.x / kRows);
  const int gkq = k0 + gk;
  const float gx = (!VEC && gkq < kpad) ? g.ct[gkq] : 0.f;
  const int sbase = VEC ? 0 : g.seg_off[m];
  const int nsg = VEC ? 0 : g.seg_off[m + 1] - sbase;
  // this order's checkpoints through a buffer descriptor: a segment past the
  // order's last, or a ring pair past the last, reads zero, with no branch
  const __amdgpu_buffer_rsrc_t ckr =
      buffer_rsrc(reinterpret_cast<const float*>(VEC ? nullptr : g.ck + (int64_t)sbase * kpad),
                  VEC ? 0 : (int64_t)nsg * kpad * 8);
  const int ckoff = gkq < kpad ? (gg * kpad + gkq) * 8 : nsg * kpad * 8;

  // table tile: this block's 64 ring pairs, i chunk starting at ic
  bool a_ok = false;  // vector transform: the last loaded table piece lies in the table block
  auto load_a = [&](int ic, float* v) {
    if constexpr (!VEC) {  // the checkpoint; store_a generates
      v[0] = buffer_load(ckr, ckoff, (ic / 8) * kpad * 8);
      v[1] = buffer_load(ckr, ckoff + 4, (ic / 8) * kpad * 8);
      return;
    }
    // an entry outside the table block reads the block's first (valid
    // memory) and is zeroed in store_a: the load itself is unconditional
    const bool half1 = VEC && ic >= kp;
    const int ii = half1 ? ic - kp : ic;
    const int spad = half1 ? spad1 : spad0;
    a_ok = k0 + sr < kpad && ii + sh < spad;
    const float* src = a_ok ? (half1 ? Pbase1 : Pbase0) + (int64_t)k0 * spad + ii : PT + p_off[m];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 t = *reinterpret_cast<const float4*>(src + 4 * q);
      v[4 * q] = t.x;
      v[4 * q + 1] = t.y;
      v[4 * q + 2] = t.z;
      v[4 * q + 3] = t.w;
    }
  };
  // the recurrence coefficients of the segment of chunk ic (wave-uniform,
  // scalar loads); a segment past the last one is never used (zero
  // checkpoint): the last segment's are read instead of running off the table
  auto load_cf = [&](int ic, float2 (&cf)[kSegL]) {
    if constexpr (!VEC) {
      const int seg = min(ic / 8 + gg, nsg - 1);
      const float2* abp = g.ab + (int64_t)(sbase + seg) * kSegL;
#pragma unroll
      for (int j = 2; j < kSegL; ++j) cf[j] = abp[j];
    }
  };
  auto store_a = [&](Lds& L, const float* v, const float2 (&cf)[kSegL]) {
    if constexpr (!VEC) {
      float w[kSegL];
      segment_recurrence(w, make_float2(v[0], v[1]), gx, [&](int j) { return cf[j]; });
      const int c8 = 8 * gg;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if constexpr (PREC == HPX_PREC_FP32) {
          *reinterpret_cast<float4*>(&L.a[p][gk][c8]) = make_float4(w[p], w[p + 2], w[p + 4], w[p + 6]);
          *reinterpret_cast<float4*>(&L.a[p][gk][c8 + 4]) = make_float4(w[p + 8], w[p + 10], w[p + 12], w[p + 14]);
        } else {
          bf16x8 hi, lo;
          const float wp[8] = {w[p], w[p + 2], w[p + 4], w[p + 6], w[p + 8], w[p + 10], w[p + 12], w[p + 14]};
          split_to<PREC>(wp, hi, lo);
          const int slot = bf16_slot(gk, c8);
          *reinterpret_cast<bf16x8*>(&L.hi[p][gk][slot]) = hi;
          if constexpr (PREC == HPX_PREC_BF16X3) *reinterpret_cast<bf16x8*>(&L.lo[p][gk][slot]) = lo;
        }
      }
      return;
    }
    float vz[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) vz[q] = a_ok ? v[q] : 0.f;
    if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.a[sp][sr][sh + q] = vz[q];
    } else if constexpr (PREC == HPX_PREC_BF16X3) {
#pragma unroll
      for (int q = 0; q < 16; ++q) split(vz[q], L.hi[sp][sr][shs + q], L.lo[sp][sr][shs + q]);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.hi[sp][sr][shs + q] = (__bf16)vz[q];
    }
  };
  // B fragments: rows i of aT (this m's rows of parity p).  Rows at or
  // beyond the parity's count (its padding, the other parity's or the next
  // m's rows, or the slack after aT) meet zero table entries; they are read
  // in bounds of the workspace and set to zero in the parity's last chunk,
  // so whatever they hold never reaches the products.
  const int kl = frag_k_lane<PREC>(lane);
  auto load_b = [&](int ic, float (&bv)[2][CT][8]) {
    const bool half1 = VEC && ic >= kp;
    const int ii = half1 ? ic - kp : ic;
    const int voff = half1 ? bvoff3 : bvoff;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int pb = VEC && !half1 ? 1 - p : p;  // coefficient parity block
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          bv[p][ct][q] = buffer_load(
              ar, voff, (int)(((int64_t)((pb ? padp[0] : 0) + ii + frag_k_uniform<PREC>(q)) * nc + ct * 16) * 4));
    }
  };
  // Just before the products (the loads have landed by then; writing a
  // register that a load in flight targets would wait for it): rows at or
  // beyond the parity's count set to zero, and the vector transform's signs.
  auto prep_b = [&](int ic, float (&bv)[2][CT][8]) {
    const bool half1 = VEC && ic >= kp;
    const int ii = half1 ? ic - kp : ic;
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int pb = VEC && !half1 ? 1 - p : p;
      if (ii + kKTile > nlp[pb]) {
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (ii + kl + frag_k_uniform<PREC>(q) >= nlp[pb]) bv[p][ct][q] = 0.f;
      }
    }
    if constexpr (VEC) {
      const float sg = half1 ? sgn1 : sgn0;
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q) bv[p][ct][q] *= sg;
    }
  };

  // Software pipeline as in the forward kernel: every half-iteration issues
  // its loads unconditionally (the last reloads the last chunk, whose tile
  // is never read); the checkpoints and the coefficients of the chunk after
  // next are requested before the products of this one.
  const int nchunk = (VEC ? 2 : 1) * kp / kKTile;  // 0 for m >= lmax: zero spectra
  if (nchunk > 0) {
    float pa[16];
    float2 cf[kSegL];
    float b0[2][CT][8], b1[2][CT][8];
    load_a(0, pa);
    load_cf(0, cf);
    load_b(0, b0);
    store_a(lds[0], pa, cf);
    __syncthreads();
    for (int c = 0;; c += 2) {
      const int c1 = min(c + 1, nchunk - 1) * kKTile;
      load_a(c1, pa);
      load_cf(c1, cf);
      load_b(c1, b1);
      prep_b(c * kKTile, b0);
      mfma_chunk<PREC, CT>(lds[0], b0, acc, lane, rh * 64);
      store_a(lds[1], pa, cf);
      __syncthreads();
      if (c + 1 >= nchunk) break;
      const int c2 = min(c + 2, nchunk - 1) * kKTile;
      load_a(c2, pa);
      load_cf(c2, cf);
      load_b(c2, b0);
      prep_b((c + 1) * kKTile, b1);
      mfma_chunk<PREC, CT>(lds[1], b1, acc, lane, rh * 64);
      store_a(lds[0], pa, cf);
      __syncthreads();
      if (c + 2 >= nchunk) break;
    }
  }

  // rows k0..k0+kRows-1 of X[m][p], this work-group's columns; none below
  // kmin[m], which no synthesis reads at this m
  const int64_t pstride = (int64_t)kpad * ncol;
  float* Xb = X + (int64_t)m * 2 * pstride + (int64_t)k0 * ncol + cb * (64 * CT);
  const int kfirst = kmin[m];
  store_tiles_lds<CT, RH, true>(acc, reinterpret_cast<float*>(smem), wave, lane, ncol - cb * (64 * CT),
                      [&](int p, int r) -> float* {
                        const int k = k0 + r;
                        return k < kpad && k >= kfirst ? Xb + p * pstride + (int64_t)r * ncol : nullptr;
                      });
}

// PT[m][p][k][i] = P[m][p][i][k]: the table in the inverse GEMM's operand
// layout (rows = ring pairs, i contiguous, padded to 16 like the rows of P),
// so that its tiles stage into LDS without a transposition.
__global__ __launch_bounds__(kThreads) void table_transpose_kernel(const float* __restrict__ P,
                                                                   float* __restrict__ PT,
                                                                   const int64_t* __restrict__ p_off,
                                                                   const int* __restrict__ nl, int kpad) {
  __shared__ float t[32][33];
  const int m = blockIdx.z >> 1;
  const int par = blockIdx.z & 1;
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  const int padp = (nl[2 * m + par] + kRowPad - 1) / kRowPad * kRowPad;
  const int i0 = blockIdx.y * 32;
  const int kk0 = blockIdx.x * 32;
  if (i0 >= padp) return;
  const int64_t base = p_off[m] + (par ? (int64_t)pad0 * kpad : 0);
  const int tx = threadIdx.x & 31;
  const int ty = threadIdx.x >> 5;
  for (int r = ty; r < 32; r += kThreads / 32) {
    const int i = i0 + r;
    t[r][tx] = i < padp ? P[base + (int64_t)i * kpad + kk0 + tx] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < 32; r += kThreads / 32) {
    const int i = i0 + tx;
    if (i < padp) PT[base + (int64_t)(kk0 + r) * padp + i] = t[tx][r];
  }
}

// ---------------------------------------------------------------------------
// Coefficient layout changes between aT (GEMM order, one row of fields per
// (m, l)) and the user layout a[b][l][m] (complex, m fastest).  A tile is one
// l, 32 consecutive m and 64 fields: 32 rows of 512 bytes on the aT side and
// 64 runs of 256 bytes on the user side, transposed through LDS.
// Tiles run m-tile fastest on one XCD (xcd_local_index), so the user-side
// lines that two neighbouring tiles share meet in that XCD's L2; degrees run
// in the order 0, lmax-1, 1, lmax-2, ... so that every XCD's contiguous share
// holds the same number of tiles below the zero triangle.  (Measured at
// nside 256, L 513, 64 fields against one l per work-group in grid order:
// fwd+inv step 637 -> 626 us; non-temporal loads or stores on either side
// made it slower.)
// ---------------------------------------------------------------------------

constexpr int kTrM = 32;        // m per tile
constexpr int kTrCols = 128;    // aT columns per tile (64 fields, re / im)

struct TrTile {
  int l, m0;
  int64_t c0;
};

// The tile of this work-group (false: beyond the last tile).
__device__ __forceinline__ bool tr_tile(int lmax, int mmax, int64_t ncol, TrTile& t) {
  const int ntm = (mmax + kTrM - 1) / kTrM;
  const int64_t nct = (ncol + kTrCols - 1) / kTrCols;
  const int64_t item = xcd_local_index(blockIdx.x, gridDim.x);
  if (item >= (int64_t)lmax * ntm * nct) return false;
  t.m0 = (int)(item % ntm) * kTrM;
  const int64_t r = item / ntm;
  t.c0 = (r % nct) * kTrCols;
  const int lp = (int)(r / nct);
  t.l = (lp & 1) ? lmax - 1 - (lp >> 1) : (lp >> 1);
  return true;
}

__device__ __forceinline__ int64_t at_row(const int64_t* p_off, const int* nl, int kpad, int m,
                                          int l) {
  const int d = l - m;
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  return p_off[m] / kpad + ((d & 1) ? pad0 : 0) + (d >> 1);
}

// a[b][l][m] from aT; zero for l < m < mmax (the reference's dense layout).
// lnorm: scale by 1 / (l (l + 1)) (1 at l = 0), the vector transform's
// normalisation (hpx_sht.py:323-328).
__global__ __launch_bounds__(kThreads) void coeff_to_user_kernel(
    const float* __restrict__ aT, float* __restrict__ alm, const int64_t* __restrict__ p_off,
    const int* __restrict__ nl, int kpad, int lmax, int mmax, int64_t ncol, int64_t nbatch, int lnorm) {
  __shared__ float t[kTrM][kTrCols + 1];
  TrTile tile;
  if (!tr_tile(lmax, mmax, ncol, tile)) return;
  const int l = tile.l, m0 = tile.m0;
  const int64_t c0 = tile.c0;
  const int mend = min(mmax, l + 1);
  const int tid = threadIdx.x;
  {
    const int r = tid >> 3;             // tile row (m - m0)
    const int cq = (tid & 7) * 4;       // float4 within a 32-float run
    const int m = m0 + r;
    if (m < mend) {
      const float* src = aT + at_row(p_off, nl, kpad, m, l) * ncol + c0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = cq + 32 * j;
        if (c0 + c < ncol) {
          const floatx4 v = *reinterpret_cast<const floatx4*>(src + c);
          t[r][c] = v.x;
          t[r][c + 1] = v.y;
          t[r][c + 2] = v.z;
          t[r][c + 3] = v.w;
        }
      }
    }
  }
  __syncthreads();
  const int mm = tid & 31;
  const int m = m0 + mm;
  if (m >= mmax) return;
  const bool zero = m >= mend;
  const float sc = (lnorm && l > 0) ? (float)(1.0 / ((double)l * (l + 1))) : 1.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int bs = (tid >> 5) + 8 * j;  // field within the tile
    const int64_t b = c0 / 2 + bs;
    if (b >= nbatch) break;
    const float2 v = zero ? make_float2(0.f, 0.f) : make_float2(sc * t[mm][2 * bs], sc * t[mm][2 * bs + 1]);
    *reinterpret_cast<float2*>(alm + 2 * ((b * lmax + l) * (int64_t)mmax + m)) = v;
  }
}

// aT from a[b][l][m] (rows m <= l only); columns of fields >= nbatch are zero.
__global__ __launch_bounds__(kThreads) void user_to_coeff_kernel(
    const float* __restrict__ alm, float* __restrict__ aT, const int64_t* __restrict__ p_off,
    const int* __restrict__ nl, int kpad, int lmax, int mmax, int64_t ncol, int64_t nbatch) {
  __shared__ float t[kTrM][kTrCols + 1];
  TrTile tile;
  if (!tr_tile(lmax, mmax, ncol, tile)) return;
  const int l = tile.l, m0 = tile.m0;
  const int64_t c0 = tile.c0;
  const int mend = min(mmax, l + 1);
  if (m0 >= mend) return;
  const int tid = threadIdx.x;
  {
    const int mm = tid & 31;
    const int m = m0 + mm;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int bs = (tid >> 5) + 8 * j;
      const int64_t b = c0 / 2 + bs;
      floatx2 v = {0.f, 0.f};
      if (m < mend && b < nbatch)
        v = *reinterpret_cast<const floatx2*>(alm + 2 * ((b * lmax + l) * (int64_t)mmax + m));
      t[mm][2 * bs] = v.x;
      t[mm][2 * bs + 1] = v.y;
    }
  }
  __syncthreads();
  const int r = tid >> 3;
  const int cq = (tid & 7) * 4;
  const int m = m0 + r;
  if (m >= mend) return;
  float* dst = aT + at_row(p_off, nl, kpad, m, l) * ncol + c0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = cq + 32 * j;
    if (c0 + c < ncol)
      *reinterpret_cast<float4*>(dst + c) = make_float4(t[r][c], t[r][c + 1], t[r][c + 2], t[r][c + 3]);
  }
}

}  // namespace

int launch_legendre_table(ShtPlan& p, const double* d_ct, hipStream_t s) {
  dim3 grid((p.kpad + kThreads - 1) / kThreads, p.mmax);
  if (p.vector) {
    vector_table_kernel<float><<<grid, kThreads, 0, s>>>(p.d_p, p.vstride, p.d_p_off, p.d_nl, d_ct, p.npair, p.kpad,
                                                  p.lmax, p.csphase, p.vnorm, p.vinverse);
  } else {
    checkpoint_kernel<<<grid, kThreads, 0, s>>>(p.d_ck, p.d_seg_off, d_ct, p.npair, p.kpad, p.lmax);
  }
  HPX_LAUNCH_CHECK();
  if (p.d_pt != nullptr && p.mmax > 0) {
    int max_pad = 0;
    for (int m = 0; m < p.mmax; ++m)
      for (int par = 0; par < 2; ++par)
        max_pad = std::max(max_pad, (p.nl[2 * m + par] + kRowPad - 1) / kRowPad * kRowPad);
    if (max_pad > 0) {
      if (2 * p.mmax > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "mmax too large");
      const dim3 tg((unsigned)(p.kpad / 32), (unsigned)((max_pad + 31) / 32), (unsigned)(2 * p.mmax));
      for (int t = 0; t < (p.vector ? 2 : 1); ++t) {
        table_transpose_kernel<<<tg, kThreads, 0, s>>>(p.d_p + t * p.vstride, p.d_pt + t * p.vstride, p.d_p_off,
                                                       p.d_nl, p.kpad);
        HPX_LAUNCH_CHECK();
      }
    }
  }
  return HPX_OK;
}

int launch_vector_table64(const ShtPlan& p, double* D, const double* d_ct, hipStream_t s) {
  dim3 grid((p.kpad + kThreads - 1) / kThreads, p.mmax);
  vector_table_kernel<double><<<grid, kThreads, 0, s>>>(D, p.vstride, p.d_p_off, p.d_nl, d_ct, p.npair, p.kpad,
                                                         p.lmax, p.csphase, p.vnorm, p.vinverse);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

// One work-group per tile (tr_tile); 0 if the tiles do not fit one grid.
static unsigned transpose_grid(const ShtPlan& p, int64_t ncol) {
  const int64_t n = (int64_t)p.lmax * ((p.mmax + kTrM - 1) / kTrM) * ((ncol + kTrCols - 1) / kTrCols);
  return n > 0x7ffffff0 ? 0u : xcd_grid(n);
}

int launch_gemm_fwd(const ShtPlan& p, int precision, int ct, const float* X, float* aT,
                    int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0 || p.n_fwd_items == 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  if (ct == kGemmFwdW8) {
    if (p.vector || p.kpad % 64 != 0)
      HPX_FAIL(HPX_ERR_UNSUPPORTED, "the 8-wave forward GEMM needs a scalar plan with ring pairs padded to 64");
    const int64_t ncb8 = (ncol + 127) / 128;
    if ((int64_t)p.n_fwd_items * ncb8 > 0x7ffffff0) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
    const dim3 grid8(xcd_grid((int64_t)p.n_fwd_items * ncb8));
    const OtfArgs otf8{p.d_ck, p.d_ab, p.d_seg_off, p.d_ct};
#define HPX_FWD8(PR)                                                                                         \
  legendre_fwd_w8_kernel<PR><<<grid8, kW8Threads, 0, s>>>(X, aT, otf8, p.d_p_off, p.d_nl, p.d_kmin, p.d_fwd_items, \
                                                         p.n_fwd_items, (int)ncb8, p.kpad, ncol)
    switch (precision) {
      case HPX_PREC_FP32: HPX_FWD8(HPX_PREC_FP32); break;
      case HPX_PREC_BF16X3: HPX_FWD8(HPX_PREC_BF16X3); break;
      case HPX_PREC_BF16: HPX_FWD8(HPX_PREC_BF16); break;
      default: HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
    }
#undef HPX_FWD8
    HPX_LAUNCH_CHECK();
    return HPX_OK;
  }
  const int64_t ncb = (ncol + 64 * ct - 1) / (64 * ct);
  if ((int64_t)p.n_fwd_items * ncb > 0x7ffffff0) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const dim3 grid(xcd_grid((int64_t)p.n_fwd_items * ncb));
  const OtfArgs otf{p.d_ck, p.d_ab, p.d_seg_off, p.d_ct};
#define HPX_FWD(PR, CT)                                                                                       \
  do {                                                                                                        \
    if (p.vector)                                                                                             \
      legendre_fwd_kernel<PR, CT, true><<<grid, kThreads, 0, s>>>(X, aT, p.d_p, otf, p.d_p_off, p.d_nl, p.d_kmin,     \
                                                                  p.d_fwd_items, p.n_fwd_items, (int)ncb, p.kpad, \
                                                                  ncol, p.vstride);                             \
    else                                                                                                      \
      legendre_fwd_kernel<PR, CT><<<grid, kThreads, 0, s>>>(X, aT, nullptr, otf, p.d_p_off, p.d_nl, p.d_kmin,   \
                                                            p.d_fwd_items, p.n_fwd_items, (int)ncb, p.kpad,   \
                                                            ncol, 0);                                         \
  } while (0)
  switch (precision * 4 + ct) {
    case HPX_PREC_FP32 * 4 + 1: HPX_FWD(HPX_PREC_FP32, 1); break;
    case HPX_PREC_FP32 * 4 + 2: HPX_FWD(HPX_PREC_FP32, 2); break;
    case HPX_PREC_BF16X3 * 4 + 1: HPX_FWD(HPX_PREC_BF16X3, 1); break;
    case HPX_PREC_BF16X3 * 4 + 2: HPX_FWD(HPX_PREC_BF16X3, 2); break;
    case HPX_PREC_BF16 * 4 + 1: HPX_FWD(HPX_PREC_BF16, 1); break;
    case HPX_PREC_BF16 * 4 + 2: HPX_FWD(HPX_PREC_BF16, 2); break;
    default: HPX_FAIL(HPX_ERR_ARG, "no forward GEMM for precision %d, %d column tiles", precision, ct);
  }
#undef HPX_FWD
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_gemm_inv(const ShtPlan& p, int precision, int ct, int rh, const float* aT, float* X,
                    int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  if (rh != 1 && rh != 2) HPX_FAIL(HPX_ERR_ARG, "no inverse GEMM with %d row halves", rh);
  if (p.vector) rh = 1;  // the vector variants are built for 64 ring pairs per work-group
  const int64_t ncol = col_pad(nbatch);
  const int64_t ncb = (ncol + 64 * ct - 1) / (64 * ct);
  const int nkb = (p.kpad + 64 * rh - 1) / (64 * rh);
  const int nitems = nkb * p.mmax;
  if ((int64_t)nitems * ncb > 0x7ffffff0) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const dim3 grid(xcd_grid((int64_t)nitems * ncb));
  const OtfArgs otf{p.d_ck, p.d_ab, p.d_seg_off, p.d_ct};
#define HPX_INV(PR, CT, RH)                                                                                    \
  do {                                                                                                         \
    if (RH == 1 && p.vector)                                                                                   \
      legendre_inv_kernel<PR, CT, 1, true><<<grid, kThreads, 0, s>>>(aT, X, p.d_pt, otf, p.d_p_off, p.d_nl, p.d_kmin, nkb, \
                                                                     nitems, (int)ncb, p.kpad, ncol, p.at_rows, \
                                                                     p.vstride);                                \
    else                                                                                                       \
      legendre_inv_kernel<PR, CT, RH><<<grid, kThreads * RH, 0, s>>>(aT, X, nullptr, otf, p.d_p_off, p.d_nl, p.d_kmin, nkb, \
                                                                     nitems, (int)ncb, p.kpad, ncol, p.at_rows, 0); \
  } while (0)
  switch (precision * 8 + ct * 2 + rh - 1) {
    case HPX_PREC_FP32 * 8 + 2: HPX_INV(HPX_PREC_FP32, 1, 1); break;
    case HPX_PREC_FP32 * 8 + 3: HPX_INV(HPX_PREC_FP32, 1, 2); break;
    case HPX_PREC_FP32 * 8 + 4: HPX_INV(HPX_PREC_FP32, 2, 1); break;
    case HPX_PREC_FP32 * 8 + 5: HPX_INV(HPX_PREC_FP32, 2, 2); break;
    case HPX_PREC_BF16X3 * 8 + 2: HPX_INV(HPX_PREC_BF16X3, 1, 1); break;
    case HPX_PREC_BF16X3 * 8 + 3: HPX_INV(HPX_PREC_BF16X3, 1, 2); break;
    case HPX_PREC_BF16X3 * 8 + 4: HPX_INV(HPX_PREC_BF16X3, 2, 1); break;
    case HPX_PREC_BF16X3 * 8 + 5: HPX_INV(HPX_PREC_BF16X3, 2, 2); break;
    case HPX_PREC_BF16 * 8 + 2: HPX_INV(HPX_PREC_BF16, 1, 1); break;
    case HPX_PREC_BF16 * 8 + 3: HPX_INV(HPX_PREC_BF16, 1, 2); break;
    case HPX_PREC_BF16 * 8 + 4: HPX_INV(HPX_PREC_BF16, 2, 1); break;
    case HPX_PREC_BF16 * 8 + 5: HPX_INV(HPX_PREC_BF16, 2, 2); break;
    default: HPX_FAIL(HPX_ERR_ARG, "no inverse GEMM for precision %d, %d column tiles", precision, ct);
  }
#undef HPX_INV
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_coeff_to_user(const ShtPlan& p, const float* aT, float* alm, int64_t nbatch,
                         hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  const unsigned grid = transpose_grid(p, ncol);
  if (grid == 0) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  coeff_to_user_kernel<<<grid, kThreads, 0, s>>>(aT, alm, p.d_p_off, p.d_nl, p.kpad, p.lmax, p.mmax, ncol,
                                                 nbatch, p.vector);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_user_to_coeff(const ShtPlan& p, const float* alm, float* aT, int64_t nbatch,
                         hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  const unsigned grid = transpose_grid(p, ncol);
  if (grid == 0) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  user_to_coeff_kernel<<<grid, kThreads, 0, s>>>(alm, aT, p.d_p_off, p.d_nl, p.kpad, p.lmax, p.mmax, ncol,
                                                 nbatch);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_legendre_fwd(const ShtPlan& p, int precision, float* ws, float* alm,
                        int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  float* aT = ws + x_floats(p, nbatch);
  const GemmTune t = p.gemm_tune(precision, col_pad(nbatch));
  int rc = launch_gemm_fwd(p, precision, t.fwd_ct, ws, aT, nbatch, s);
  if (rc != HPX_OK) return rc;
  return launch_coeff_to_user(p, aT, alm, nbatch, s);
}

int launch_legendre_inv(const ShtPlan& p, int precision, const float* alm, float* ws,
                        int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  if (precision < 0 || precision > 2) HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
  float* aT = ws + x_floats(p, nbatch);
  const GemmTune t = p.gemm_tune(precision, col_pad(nbatch));
  int rc = launch_user_to_coeff(p, alm, aT, nbatch, s);
  if (rc != HPX_OK) return rc;
  return launch_gemm_inv(p, precision, t.inv_ct, t.inv_rh, aT, ws, nbatch, s);
}

}  // namespace hpxmi
