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
//  * Forward work-group: one m, 64 rows of each parity, 128 columns
//    (64 fields, re/im).  The table rows are staged once through LDS and
//    shared by 4 waves; each wave streams its 32 columns of the spectra
//    straight into MFMA B registers.  The output a[b][l][m] is m-minor, so the
//    work list puts consecutive m of the same row block on the same XCD,
//    whose L2 merges the partial-line writes.
//  * Inverse work-group: one m, 64 ring pairs, 128 columns; the table tile is
//    transposed through LDS.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

namespace hpxmi {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kRowsPerParity = 64;  // forward: rows i per parity per work-group
constexpr int kKBlock = 64;         // inverse: ring pairs per work-group
constexpr int kLdsPadF = 1;         // FP32 LDS row padding (floats)
constexpr int kLdsPadH = 8;         // BF16 LDS row padding (bf16)

__device__ __forceinline__ void split(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// ---------------------------------------------------------------------------
// Table generation (FP64, scaled recurrence)
// ---------------------------------------------------------------------------

// One thread per (m, k): P_mm by the sectoral product, then the three-term
// recurrence in l (cuhpx/sht_tools.py:243-252), each value carried as
// v * 2^(600 e); a value is stored when e == 0 and as 0 while e < 0
// (|P| < 2^-600, far below FP32's range).
__global__ __launch_bounds__(kThreads) void legendre_table_kernel(
    float* __restrict__ P, const int64_t* __restrict__ p_off, const int* __restrict__ nl,
    const double* __restrict__ ct, int npair, int kpad, int lmax) {
  const int k = blockIdx.x * kThreads + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= kpad) return;
  const int nl0 = nl[2 * m], nl1 = nl[2 * m + 1];
  const int pad0 = (nl0 + kRowPad - 1) / kRowPad * kRowPad;
  const int pad1 = (nl1 + kRowPad - 1) / kRowPad * kRowPad;
  float* P0 = P + p_off[m];
  float* P1 = P0 + (int64_t)pad0 * kpad;
  // padded rows and padded ring pairs are zero
  for (int i = nl0; i < pad0; ++i) P0[(int64_t)i * kpad + k] = 0.f;
  for (int i = nl1; i < pad1; ++i) P1[(int64_t)i * kpad + k] = 0.f;
  if (k >= npair) {
    for (int i = 0; i < nl0; ++i) P0[(int64_t)i * kpad + k] = 0.f;
    for (int i = 0; i < nl1; ++i) P1[(int64_t)i * kpad + k] = 0.f;
    return;
  }
  if (m >= lmax) return;
  const double x = ct[k];
  const double s = sqrt((1.0 - x) * (1.0 + x));
  const double big = 0x1p600, tiny = 0x1p-600;
  double pmm = 0.28209479177387814347;  // 1 / sqrt(4 pi)
  int e = 0;
  for (int i = 1; i <= m; ++i) {
    pmm *= sqrt((2.0 * i + 1.0) / (2.0 * i)) * s;
    if (pmm != 0.0 && fabs(pmm) < tiny) {
      pmm *= big;
      --e;
    }
  }
  // l = m
  P0[k] = e == 0 ? (float)pmm : 0.f;
  if (m + 1 >= lmax) return;
  // l = m + 1
  double p2 = pmm;                              // P_{l-2}
  double p1 = sqrt(2.0 * m + 3.0) * x * pmm;    // P_{l-1}
  if (e < 0 && fabs(p1) > 1.0) {
    p1 *= tiny;
    p2 *= tiny;
    ++e;
  }
  P1[k] = e == 0 ? (float)p1 : 0.f;
  const double dm = m;
  for (int l = m + 2; l < lmax; ++l) {
    const double dl = l;
    const double a = sqrt((2 * dl - 1) / (dl - dm) * (2 * dl + 1) / (dl + dm));
    const double b =
        sqrt((dl + dm - 1) / (dl - dm) * (2 * dl + 1) / (2 * dl - 3) * (dl - dm - 1) / (dl + dm));
    const double v = x * a * p1 - b * p2;
    p2 = p1;
    p1 = v;
    if (e < 0 && fabs(p1) > 1.0) {
      p1 *= tiny;
      p2 *= tiny;
      ++e;
    }
    float* dst = ((l - m) & 1) ? P1 : P0;
    dst[(int64_t)((l - m) >> 1) * kpad + k] = e == 0 ? (float)p1 : 0.f;
  }
}

// ---------------------------------------------------------------------------
// Forward: aT[m][p][i][col] = sum_k P[m][p][i][k] X[m][p][k][col]
// ---------------------------------------------------------------------------

template <int PREC>
struct FwdLds;

template <>
struct FwdLds<HPX_PREC_FP32> {
  float a[2][kRowsPerParity][kKTile + kLdsPadF];
};

template <>
struct FwdLds<HPX_PREC_BF16X3> {
  __bf16 hi[2][kRowsPerParity][kKTile + kLdsPadH];
  __bf16 lo[2][kRowsPerParity][kKTile + kLdsPadH];
};

template <>
struct FwdLds<HPX_PREC_BF16> {
  __bf16 hi[2][kRowsPerParity][kKTile + kLdsPadH];
};

template <int PREC, int CT>
__global__ __launch_bounds__(kThreads) void legendre_fwd_kernel(
    const float* __restrict__ X, float* __restrict__ aT, const float* __restrict__ P,
    const int64_t* __restrict__ p_off, const int* __restrict__ nl,
    const int2* __restrict__ items, int nitems, int npair, int kpad, int64_t ncol) {
  __shared__ FwdLds<PREC> lds[2];
  const int item = xcd_local_index(blockIdx.x, gridDim.x);
  if (item >= nitems) return;
  const int m = items[item].x;
  const int jb = items[item].y;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col0 = blockIdx.y * (64 * CT) + wave * (16 * CT);
  const bool wave_active = col0 < ncol;
  const int nlp[2] = {nl[2 * m], nl[2 * m + 1]};
  const int padp0 = (nlp[0] + kRowPad - 1) / kRowPad * kRowPad;
  const int padp1 = (nlp[1] + kRowPad - 1) / kRowPad * kRowPad;
  const int row0 = jb * kRowsPerParity;
  const float* Pm = P + p_off[m];
  const int64_t pstride = (int64_t)kpad * ncol;
  const float* Xm = X + (int64_t)m * 2 * pstride;

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
  const int srow = row0 + sr;
  const bool srow_ok = srow < (sp ? padp1 : padp0);
  const float* Psrc = Pm + (sp ? (int64_t)padp0 * kpad : 0) + (int64_t)srow * kpad + sh;

  // table tile (16 floats of one row) -> registers
  auto load_a = [&](int kc, float* v) {
    if (srow_ok) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 t = *reinterpret_cast<const float4*>(Psrc + kc + 4 * q);
        v[4 * q] = t.x;
        v[4 * q + 1] = t.y;
        v[4 * q + 2] = t.z;
        v[4 * q + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = 0.f;
    }
  };
  // registers -> LDS (split into BF16 hi / lo here for the BF16 paths)
  auto store_a = [&](FwdLds<PREC>& L, const float* v) {
    if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.a[sp][sr][sh + q] = v[q];
    } else if constexpr (PREC == HPX_PREC_BF16X3) {
#pragma unroll
      for (int q = 0; q < 16; ++q) split(v[q], L.hi[sp][sr][sh + q], L.lo[sp][sr][sh + q]);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.hi[sp][sr][sh + q] = (__bf16)v[q];
    }
  };
  // this wave's B operand (spectra) for one k chunk, in MFMA lane order
  auto load_b = [&](int kc, float (&bv)[2][CT][8]) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int col = col0 + ct * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          int kk;
          if constexpr (PREC == HPX_PREC_FP32)
            kk = kc + 4 * q + (lane >> 4);
          else
            kk = kc + 8 * (lane >> 4) + q;
          bv[p][ct][q] = kk < npair ? Xm[p * pstride + (int64_t)kk * ncol + col] : 0.f;
        }
      }
  };
  auto compute = [&](const FwdLds<PREC>& L, const float (&bv)[2][CT][8]) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          if (row0 + rt * 16 >= nlp[p]) continue;
          const int ar = rt * 16 + (lane & 15);
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float av = L.a[p][ar][4 * q + (lane >> 4)];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
              acc[p][rt][ct] =
                  __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[p][ct][q], acc[p][rt][ct], 0, 0, 0);
          }
        }
      } else {
        bf16x8 bhi[CT], blo[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            __bf16 h, l;
            split(bv[p][ct][q], h, l);
            bhi[ct][q] = h;
            blo[ct][q] = l;
          }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          if (row0 + rt * 16 >= nlp[p]) continue;
          const int ar = rt * 16 + (lane & 15);
          const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(&L.hi[p][ar][8 * (lane >> 4)]);
          bf16x8 alo;
          if constexpr (PREC == HPX_PREC_BF16X3)
            alo = *reinterpret_cast<const bf16x8*>(&L.lo[p][ar][8 * (lane >> 4)]);
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
  };

  // software pipeline: chunk c+1's loads are in flight while chunk c is
  // multiplied; the table tile is double-buffered in LDS (one barrier per chunk)
  const int nchunk = kpad / kKTile;
  float pa[16];
  float bv[2][CT][8];
  load_a(0, pa);
  if (wave_active) load_b(0, bv);
  store_a(lds[0], pa);
  __syncthreads();
  for (int c = 0; c < nchunk; ++c) {
    const bool more = c + 1 < nchunk;
    float pn[16];
    float bn[2][CT][8];
    if (more) {
      load_a((c + 1) * kKTile, pn);
      if (wave_active) load_b((c + 1) * kKTile, bn);
    }
    if (wave_active) compute(lds[c & 1], bv);
    if (more) {
      store_a(lds[(c + 1) & 1], pn);
#pragma unroll
      for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q) bv[p][ct][q] = bn[p][ct][q];
    }
    __syncthreads();
  }

  if (!wave_active) return;
  // ---- write this m's rows of aT: 16 lanes store 64 contiguous bytes of a row
  float* Am = aT + (p_off[m] / kpad) * ncol;
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int col = col0 + ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = row0 + rt * 16 + 4 * (lane >> 4) + r;
          if (i < nlp[p]) Am[(int64_t)((p ? padp0 : 0) + i) * ncol + col] = acc[p][rt][ct][r];
        }
      }
}

// ---------------------------------------------------------------------------
// Inverse: Y[m][p][k][col] = sum_i P[m][p][i][k] aT[m][p][i][col]
// ---------------------------------------------------------------------------

template <int PREC>
struct InvLds;

template <>
struct InvLds<HPX_PREC_FP32> {
  float a[2][kKBlock][kKTile + kLdsPadF];
};

template <>
struct InvLds<HPX_PREC_BF16X3> {
  __bf16 hi[2][kKBlock][kKTile + kLdsPadH];
  __bf16 lo[2][kKBlock][kKTile + kLdsPadH];
};

template <>
struct InvLds<HPX_PREC_BF16> {
  __bf16 hi[2][kKBlock][kKTile + kLdsPadH];
};

template <int PREC, int CT>
__global__ __launch_bounds__(kThreads) void legendre_inv_kernel(
    const float* __restrict__ aT, float* __restrict__ X, const float* __restrict__ P,
    const int64_t* __restrict__ p_off, const int* __restrict__ nl, int nkb, int nitems, int kpad,
    int64_t ncol) {
  __shared__ InvLds<PREC> lds[2];
  // m outer, ring-pair block inner: the nkb work-groups of one m read the same
  // coefficient rows aT[m] and land on one XCD (contiguous item ranges)
  const int item = xcd_local_index(blockIdx.x, gridDim.x);
  if (item >= nitems) return;
  const int m = item / nkb;
  const int kb = item - m * nkb;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col0 = blockIdx.y * (64 * CT) + wave * (16 * CT);
  const bool wave_active = col0 < ncol;
  const int nlp[2] = {nl[2 * m], nl[2 * m + 1]};
  const int padp0 = (nlp[0] + kRowPad - 1) / kRowPad * kRowPad;
  const int padp1 = (nlp[1] + kRowPad - 1) / kRowPad * kRowPad;
  const int nmax = nlp[0] > nlp[1] ? nlp[0] : nlp[1];
  const float* Pm = P + p_off[m];
  const int64_t pstride = (int64_t)kpad * ncol;
  const int k0 = kb * kKBlock;

  floatx4 acc[2][4][CT];
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) acc[p][rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};

  // staging: 2 parities x 32 rows i x 4 quarters of 16 ring pairs
  const int sp = threadIdx.x >> 7;
  const int si = (threadIdx.x >> 2) & 31;
  const int sq = (threadIdx.x & 3) * 16;
  const int spad = sp ? padp1 : padp0;
  const float* Pbase = Pm + (sp ? (int64_t)padp0 * kpad : 0);

  auto load_a = [&](int ic, float* v) {
    const int irow = ic + si;
    const int kk = k0 + sq;
    if (irow < spad && kk < kpad) {
      const float* src = Pbase + (int64_t)irow * kpad + kk;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 t = *reinterpret_cast<const float4*>(src + 4 * q);
        v[4 * q] = t.x;
        v[4 * q + 1] = t.y;
        v[4 * q + 2] = t.z;
        v[4 * q + 3] = t.w;
      }
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) v[q] = 0.f;
    }
  };
  auto store_a = [&](InvLds<PREC>& L, const float* v) {
    if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.a[sp][sq + q][si] = v[q];
    } else if constexpr (PREC == HPX_PREC_BF16X3) {
#pragma unroll
      for (int q = 0; q < 16; ++q) split(v[q], L.hi[sp][sq + q][si], L.lo[sp][sq + q][si]);
    } else {
#pragma unroll
      for (int q = 0; q < 16; ++q) L.hi[sp][sq + q][si] = (__bf16)v[q];
    }
  };
  const float* Am = aT + (p_off[m] / kpad) * ncol;
  auto load_b = [&](int ic, float (&bv)[2][CT][8]) {
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int col = col0 + ct * 16 + (lane & 15);
        const float* src = Am + (int64_t)(p ? padp0 : 0) * ncol + col;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          int i;
          if constexpr (PREC == HPX_PREC_FP32)
            i = ic + 4 * q + (lane >> 4);
          else
            i = ic + 8 * (lane >> 4) + q;
          bv[p][ct][q] = i < nlp[p] ? src[(int64_t)i * ncol] : 0.f;
        }
      }
  };
  auto compute = [&](const InvLds<PREC>& L, const float (&bv)[2][CT][8], int ic) {
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      if (ic >= nlp[p]) continue;
      if constexpr (PREC == HPX_PREC_FP32) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          const int ar = rt * 16 + (lane & 15);
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float av = L.a[p][ar][4 * q + (lane >> 4)];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct)
              acc[p][rt][ct] =
                  __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[p][ct][q], acc[p][rt][ct], 0, 0, 0);
          }
        }
      } else {
        bf16x8 bhi[CT], blo[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            __bf16 h, l;
            split(bv[p][ct][q], h, l);
            bhi[ct][q] = h;
            blo[ct][q] = l;
          }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          const int ar = rt * 16 + (lane & 15);
          const bf16x8 ahi = *reinterpret_cast<const bf16x8*>(&L.hi[p][ar][8 * (lane >> 4)]);
          bf16x8 alo;
          if constexpr (PREC == HPX_PREC_BF16X3)
            alo = *reinterpret_cast<const bf16x8*>(&L.lo[p][ar][8 * (lane >> 4)]);
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
  };

  const int nchunk = (nmax + kKTile - 1) / kKTile;
  if (nchunk > 0) {
    float pa[16];
    float bv[2][CT][8];
    load_a(0, pa);
    if (wave_active) load_b(0, bv);
    store_a(lds[0], pa);
    __syncthreads();
    for (int c = 0; c < nchunk; ++c) {
      const bool more = c + 1 < nchunk;
      float pn[16];
      float bn[2][CT][8];
      if (more) {
        load_a((c + 1) * kKTile, pn);
        if (wave_active) load_b((c + 1) * kKTile, bn);
      }
      if (wave_active) compute(lds[c & 1], bv, c * kKTile);
      if (more) {
        store_a(lds[(c + 1) & 1], pn);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
          for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int q = 0; q < 8; ++q) bv[p][ct][q] = bn[p][ct][q];
      }
      __syncthreads();
    }
  }

  if (!wave_active) return;
  float* Xm = X + (int64_t)m * 2 * pstride;
#pragma unroll
  for (int p = 0; p < 2; ++p)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < CT; ++ct) {
        const int col = col0 + ct * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = k0 + rt * 16 + 4 * (lane >> 4) + r;
          if (k < kpad) Xm[p * pstride + (int64_t)k * ncol + col] = acc[p][rt][ct][r];
        }
      }
}

// ---------------------------------------------------------------------------
// Coefficient layout changes between aT (GEMM order, one row of fields per
// (m, l)) and the user layout a[b][l][m] (complex, m fastest).  A tile is one
// l, 32 consecutive m and 64 fields: 32 rows of 512 bytes on the aT side and
// 64 runs of 256 bytes on the user side, transposed through LDS.
// ---------------------------------------------------------------------------

constexpr int kTrM = 32;        // m per tile
constexpr int kTrCols = 128;    // aT columns per tile (64 fields, re / im)

__device__ __forceinline__ int64_t at_row(const int64_t* p_off, const int* nl, int kpad, int m,
                                          int l) {
  const int d = l - m;
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  return p_off[m] / kpad + ((d & 1) ? pad0 : 0) + (d >> 1);
}

// a[b][l][m] from aT; zero for l < m < mmax (the reference's dense layout).
__global__ __launch_bounds__(kThreads) void coeff_to_user_kernel(
    const float* __restrict__ aT, float* __restrict__ alm, const int64_t* __restrict__ p_off,
    const int* __restrict__ nl, int kpad, int lmax, int mmax, int64_t ncol, int64_t nbatch) {
  __shared__ float t[kTrM][kTrCols + 1];
  const int l = blockIdx.x;
  const int m0 = blockIdx.y * kTrM;
  const int64_t c0 = (int64_t)blockIdx.z * kTrCols;
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
          const float4 v = *reinterpret_cast<const float4*>(src + c);
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
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int bs = (tid >> 5) + 8 * j;  // field within the tile
    const int64_t b = c0 / 2 + bs;
    if (b >= nbatch) break;
    const float2 v = zero ? make_float2(0.f, 0.f) : make_float2(t[mm][2 * bs], t[mm][2 * bs + 1]);
    *reinterpret_cast<float2*>(alm + 2 * ((b * lmax + l) * (int64_t)mmax + m)) = v;
  }
}

// aT from a[b][l][m] (rows m <= l only); columns of fields >= nbatch are zero.
__global__ __launch_bounds__(kThreads) void user_to_coeff_kernel(
    const float* __restrict__ alm, float* __restrict__ aT, const int64_t* __restrict__ p_off,
    const int* __restrict__ nl, int kpad, int lmax, int mmax, int64_t ncol, int64_t nbatch) {
  __shared__ float t[kTrM][kTrCols + 1];
  const int l = blockIdx.x;
  const int m0 = blockIdx.y * kTrM;
  const int64_t c0 = (int64_t)blockIdx.z * kTrCols;
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
      float2 v = make_float2(0.f, 0.f);
      if (m < mend && b < nbatch)
        v = *reinterpret_cast<const float2*>(alm + 2 * ((b * lmax + l) * (int64_t)mmax + m));
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
  legendre_table_kernel<<<grid, kThreads, 0, s>>>(p.d_p, p.d_p_off, p.d_nl, d_ct, p.npair, p.kpad,
                                                  p.lmax);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

static dim3 transpose_grid(const ShtPlan& p, int64_t ncol) {
  return dim3((unsigned)p.lmax, (unsigned)((p.mmax + kTrM - 1) / kTrM),
              (unsigned)((ncol + kTrCols - 1) / kTrCols));
}

int launch_gemm_fwd(const ShtPlan& p, int precision, int ct, const float* X, float* aT,
                    int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0 || p.n_fwd_items == 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  const int64_t ycols = (ncol + 64 * ct - 1) / (64 * ct);
  if (ycols > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const dim3 grid(xcd_grid(p.n_fwd_items), (unsigned)ycols);
#define HPX_FWD(PR, CT)                                                                      \
  legendre_fwd_kernel<PR, CT><<<grid, kThreads, 0, s>>>(X, aT, p.d_p, p.d_p_off, p.d_nl,     \
                                                        p.d_fwd_items, p.n_fwd_items, p.npair, \
                                                        p.kpad, ncol)
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

int launch_gemm_inv(const ShtPlan& p, int precision, int ct, const float* aT, float* X,
                    int64_t nbatch, hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  const int64_t ycols = (ncol + 64 * ct - 1) / (64 * ct);
  if (ycols > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const int nkb = (p.kpad + kKBlock - 1) / kKBlock;
  const int nitems = nkb * p.mmax;
  const dim3 grid(xcd_grid(nitems), (unsigned)ycols);
#define HPX_INV(PR, CT)                                                                           \
  legendre_inv_kernel<PR, CT><<<grid, kThreads, 0, s>>>(aT, X, p.d_p, p.d_p_off, p.d_nl, nkb, nitems, \
                                                        p.kpad, ncol)
  switch (precision * 4 + ct) {
    case HPX_PREC_FP32 * 4 + 1: HPX_INV(HPX_PREC_FP32, 1); break;
    case HPX_PREC_FP32 * 4 + 2: HPX_INV(HPX_PREC_FP32, 2); break;
    case HPX_PREC_BF16X3 * 4 + 1: HPX_INV(HPX_PREC_BF16X3, 1); break;
    case HPX_PREC_BF16X3 * 4 + 2: HPX_INV(HPX_PREC_BF16X3, 2); break;
    case HPX_PREC_BF16 * 4 + 1: HPX_INV(HPX_PREC_BF16, 1); break;
    case HPX_PREC_BF16 * 4 + 2: HPX_INV(HPX_PREC_BF16, 2); break;
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
  if ((ncol + kTrCols - 1) / kTrCols > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  coeff_to_user_kernel<<<transpose_grid(p, ncol), kThreads, 0, s>>>(aT, alm, p.d_p_off, p.d_nl, p.kpad,
                                                                   p.lmax, p.mmax, ncol, nbatch);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_user_to_coeff(const ShtPlan& p, const float* alm, float* aT, int64_t nbatch,
                         hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ncol = col_pad(nbatch);
  if ((ncol + kTrCols - 1) / kTrCols > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  user_to_coeff_kernel<<<transpose_grid(p, ncol), kThreads, 0, s>>>(alm, aT, p.d_p_off, p.d_nl, p.kpad,
                                                                   p.lmax, p.mmax, ncol, nbatch);
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
  return launch_gemm_inv(p, precision, t.inv_ct, aT, ws, nbatch, s);
}

}  // namespace hpxmi
