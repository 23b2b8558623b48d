// FP64 spherical harmonic transforms: the float64 -> complex128 contract of
// the drop-in.  The reference dispatches float64 maps to complex128 spectra
// (src/harmonic_transform/hpx_fft.cpp:45: kDouble -> kComplexDouble) and
// contracts them with the FP64 table cast to the input's dtype
// (cuhpx/hpx_sht.py:557, 599: weights.to(x.dtype)), so SHTCUDA on float64
// maps computes in FP64 throughout.
//
// The same operator as the FP32 path, in FP64 end to end:
//  * per ring pair and map, the packed DFT of the two mirror rings
//    (z = f_t + i f_t') by a work-group-wide radix-4/2 Stockham FFT in LDS,
//    Bluestein's chirp-z algorithm for lengths that are not a power of two
//    (tables built in FP64 on the host), then the folded spectra
//    S = w (F_t + F_t'), D = w (F_t - F_t') in the layout of the FP32 path;
//  * the Legendre table in FP64 (generated on the device by the scaled
//    three-term recurrence, kept in the plan after hpx_sht_plan_enable_f64),
//    and LDS-tiled FP64 GEMMs per (m, parity) on the folded operands;
//  * the coefficient layout changes between GEMM order and a[b][l][m].
// Launch shapes are fixed: the FP64 path is the accuracy contract, not the
// throughput path.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

#include <cmath>
#include <complex>
#include <mutex>
#include <vector>

namespace hpxmi {
namespace {

using cd = std::complex<double>;

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 zconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 zscale(double2 a, double s) { return make_double2(s * a.x, s * a.y); }

constexpr int kT64 = 512;      // threads of an FFT work-group
constexpr int kMaxM64 = 8192;  // longest FFT (128 KB of LDS)
constexpr int kGemmThreads = 256;

// ---------------------------------------------------------------------------
// Legendre table in FP64 (layout of the FP32 table: P[m][parity][i][kpad],
// rows padded to kRowPad, zero beyond the last row and ring pair).
// One thread per (m, k): P_mm by the sectoral product, then the three-term
// recurrence in l (cuhpx/sht_tools.py:243-252), each value carried as
// v * 2^(600 e) and stored when e == 0 (0 while e < 0: |P| < 2^-600).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void legendre_table64_kernel(double* __restrict__ P,
                                                               const int64_t* __restrict__ p_off,
                                                               const int* __restrict__ nl,
                                                               const double* __restrict__ ct, int npair,
                                                               int kpad, int lmax) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const int m = blockIdx.y;
  if (k >= kpad) return;
  const int nl0 = nl[2 * m], nl1 = nl[2 * m + 1];
  const int pad0 = (nl0 + kRowPad - 1) / kRowPad * kRowPad;
  const int pad1 = (nl1 + kRowPad - 1) / kRowPad * kRowPad;
  double* P0 = P + p_off[m];
  double* P1 = P0 + (int64_t)pad0 * kpad;
  for (int i = nl0; i < pad0; ++i) P0[(int64_t)i * kpad + k] = 0.0;
  for (int i = nl1; i < pad1; ++i) P1[(int64_t)i * kpad + k] = 0.0;
  if (k >= npair) {
    for (int i = 0; i < nl0; ++i) P0[(int64_t)i * kpad + k] = 0.0;
    for (int i = 0; i < nl1; ++i) P1[(int64_t)i * kpad + k] = 0.0;
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
  P0[k] = e == 0 ? pmm : 0.0;
  if (m + 1 >= lmax) return;
  double p2 = pmm;
  double p1 = sqrt(2.0 * m + 3.0) * x * pmm;
  if (e < 0 && fabs(p1) > 1.0) {
    p1 *= tiny;
    p2 *= tiny;
    ++e;
  }
  P1[k] = e == 0 ? p1 : 0.0;
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
    double* dst = ((l - m) & 1) ? P1 : P0;
    dst[(int64_t)((l - m) >> 1) * kpad + k] = e == 0 ? p1 : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Work-group-wide FFT in LDS (natural order in and out, in place): Stockham
// passes of radix 4 (radix 2 once for an odd log2 M); every thread holds its
// butterflies' inputs in registers across the barrier, then writes the outputs.
// ---------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void dft64(double2* v) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = zadd(a, b);
    v[1] = zsub(a, b);
  } else {
    const double2 t0 = zadd(v[0], v[2]), t1 = zsub(v[0], v[2]);
    const double2 t2 = zadd(v[1], v[3]);
    const double2 d = zsub(v[1], v[3]);
    const double2 t3 = make_double2(d.y, -d.x);  // -i (v1 - v3)
    v[0] = zadd(t0, t2);
    v[1] = zadd(t1, t3);
    v[2] = zsub(t0, t2);
    v[3] = zsub(t1, t3);
  }
}

template <int R>
__device__ void pass64(double2* buf, int log2M, int Ns, int log2Ns, const double2* __restrict__ tw,
                       int log2_tab) {
  constexpr int kLogR = R == 4 ? 2 : 1;
  constexpr int kMaxB = (kMaxM64 / R + kT64 - 1) / kT64;
  const int log2mr = log2M - kLogR;
  const int nb = 1 << log2mr;
  const int tw_shift = log2_tab - log2Ns - kLogR;
  double2 v[kMaxB][R];
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kT64;
    if (b < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[b + (r << log2mr)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kT64;
    if (b < nb) {
      const int jn = b & (Ns - 1);
      if (Ns > 1) {
        const double2 w = tw[jn << tw_shift];
        double2 wr = w;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[i][r] = zmul(v[i][r], wr);
          if (r + 1 < R) wr = zmul(wr, w);
        }
      }
      dft64<R>(v[i]);
      double2* dst = buf + (b - jn) * R + jn;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

__device__ void fft64(double2* buf, int log2M, const double2* __restrict__ tw, int log2_tab) {
  int Ns = 1, log2Ns = 0, rem = log2M;
  while (rem > 0) {
    if (rem == 1) {
      pass64<2>(buf, log2M, Ns, log2Ns, tw, log2_tab);
      rem -= 1;
      Ns <<= 1;
      log2Ns += 1;
    } else {
      pass64<4>(buf, log2M, Ns, log2Ns, tw, log2_tab);
      rem -= 2;
      Ns <<= 2;
      log2Ns += 2;
    }
  }
}

// DFT of length N = pr.nseq of the sequence in buf (points x < N; zero beyond): a
// power-of-two FFT, or Bluestein: X[x] = chirp_x conj(c_x),
// c = FFT(conj(FFT(z chirp) bhat)), bhat = FFT(conj chirp) / M.
__device__ void ring_dft64(double2* buf, const RingPair& pr, const double2* __restrict__ chirp,
                           const double2* __restrict__ bhat, const double2* __restrict__ tw, int log2_tab) {
  const int N = pr.nseq, M = pr.M;
  if (M == N) {
    fft64(buf, pr.log2M, tw, log2_tab);
    return;
  }
  for (int x = threadIdx.x; x < M; x += kT64) buf[x] = x < N ? zmul(buf[x], chirp[x]) : make_double2(0.0, 0.0);
  __syncthreads();
  fft64(buf, pr.log2M, tw, log2_tab);
  for (int x = threadIdx.x; x < M; x += kT64) buf[x] = zconj(zmul(buf[x], bhat[x]));
  __syncthreads();
  fft64(buf, pr.log2M, tw, log2_tab);
  for (int x = threadIdx.x; x < N; x += kT64) buf[x] = zmul(chirp[x], zconj(buf[x]));
  __syncthreads();
}

struct Args64 {
  const RingPair* pairs;
  const double2* phase;   // [npair][mmax] exp(-i m phi0)
  const double2* chirp;   // indexed by pairs[k].chirp_off
  const double2* bhat;    // indexed by pairs[k].bhat_off
  const double2* tw;      // [max_M] exp(-2 pi i x / max_M)
  const double* w;        // [npair] ring weights
  int log2_tab;
  int64_t npix, nbatch, ncol;
  int mmax, kpad, nring;
  int use_weight, cmode, layout;
  double scale;
};

// Rings transformed one at a time (RingPair::half, set for the FP64 FFT
// length limit): the FP32 kernels' ring-by-ring forms in FP64.  The ring's
// N samples are H = N/2 complex points z_j = f_2j + i f_2j+1 with
//   F[m] = E[m] + e^{-2 pi i m / N} O[m],  E = (Z[m] + conj Z[H-m]) / 2,
//   O = (Z[m] - conj Z[H-m]) / (2i),  Z = DFT_H(z)  (indices mod H),
// inversely f_2n + i f_2n+1 = conj(DFT_H(conj Y))[n], Y = E' + i O',
// E' = F[m] + F[m+H], O' = (F[m] - F[m+H]) e^{2 pi i m / N}; half = 2 takes
// DFT_H as A[x mod Q] + e^{-2 pi i x / H} B[x mod Q] over the Q = H/2-point
// DFTs of z's even and odd points (A kept in registers).
constexpr int kHalfRegs64 = kMaxM64 / 2 / kT64;

template <class Load>
__device__ __forceinline__ void ring_seq_dft64(double2* buf, const Args64& a, const RingPair& pr, Load load) {
  for (int j = threadIdx.x; j < pr.M; j += kT64) buf[j] = j < pr.nseq ? load(j) : make_double2(0.0, 0.0);
  __syncthreads();
  ring_dft64(buf, pr, a.chirp + pr.chirp_off, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
}

__device__ void analysis_half64(double2* buf, const Args64& a, const RingPair& pr, int k, int64_t b,
                                const double* __restrict__ fr, double* __restrict__ X) {
  const int N = pr.nphi, H = N >> 1, Q = pr.nseq;
  const double wscale = (a.use_weight ? a.w[k] : 1.0) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const double2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int side = 0; side < (pr.single ? 1 : 2); ++side) {
    const double* src = fr + (side ? pr.off_s : pr.off_n);
    if (pr.half == 1) {
      ring_seq_dft64(buf, a, pr, [&](int j) { return make_double2(src[2 * j], src[2 * j + 1]); });
    } else {
      double2 Areg[kHalfRegs64];
      ring_seq_dft64(buf, a, pr, [&](int j) { return make_double2(src[4 * j], src[4 * j + 1]); });
#pragma unroll
      for (int i = 0; i < kHalfRegs64; ++i) {
        const int x = threadIdx.x + i * kT64;
        if (x < Q) Areg[i] = buf[x];
      }
      __syncthreads();
      ring_seq_dft64(buf, a, pr, [&](int j) { return make_double2(src[4 * j + 2], src[4 * j + 3]); });
#pragma unroll
      for (int i = 0; i < kHalfRegs64; ++i) {
        const int x = threadIdx.x + i * kT64;
        if (x < Q) buf[Q + x] = Areg[i];
      }
      __syncthreads();
    }
    auto zval = [&](int x) -> double2 {
      if (pr.half == 1) return buf[x];
      const int xq = x < Q ? x : x - Q;
      double sn, cs;
      sincospi(-2.0 * x / H, &sn, &cs);
      return zadd(buf[Q + xq], zmul(make_double2(cs, sn), buf[xq]));
    };
    for (int m = threadIdx.x; m < a.mmax; m += kT64) {
      double2 F = make_double2(0.0, 0.0);
      if (m < K) {
        const double2 g1 = zval(m == H ? 0 : m);
        const double2 g2 = zval(m == 0 ? 0 : H - m);
        const double2 E = make_double2(0.5 * (g1.x + g2.x), 0.5 * (g1.y - g2.y));
        const double2 O = make_double2(0.5 * (g1.y + g2.y), -0.5 * (g1.x - g2.x));
        double sn, cs;
        sincospi(-2.0 * m / N, &sn, &cs);
        F = zadd(E, zmul(make_double2(cs, sn), O));
        const double c = (a.cmode == 1 && m != 0 && 2 * m != N) ? 2.0 * wscale : wscale;
        F = zscale(zmul(F, ph[m]), c);
      }
      if (a.layout == 0) {
        double2* S = reinterpret_cast<double2*>(X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b);
        double2* D = reinterpret_cast<double2*>(reinterpret_cast<double*>(S) + pstride);
        if (side == 0) {
          *S = F;
          *D = F;
        } else {
          const double2 Fn = *S;
          *S = zadd(Fn, F);
          *D = zsub(Fn, F);
        }
      } else {
        double2* out = reinterpret_cast<double2*>(X) + b * a.nring * (int64_t)a.mmax;
        out[(int64_t)(side ? a.nring - 1 - k : k) * a.mmax + m] = F;
      }
    }
    __syncthreads();
  }
}

__device__ void synthesis_half64(double2* buf, const Args64& a, const RingPair& pr, int k, int64_t b,
                                 const double* __restrict__ X, double* __restrict__ fr) {
  const int N = pr.nphi, H = N >> 1, Q = pr.nseq;
  const double wscale = (a.use_weight ? a.w[k] : 1.0) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const double2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int side = 0; side < (pr.single ? 1 : 2); ++side) {
    auto spec = [&](int m) -> double2 {
      if (m >= K) return make_double2(0.0, 0.0);
      double2 G;
      if (a.layout == 0) {
        const double* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
        const double2 S = *reinterpret_cast<const double2*>(base);
        const double2 D = *reinterpret_cast<const double2*>(base + pstride);
        G = (side == 0 || pr.single) ? zadd(S, D) : zsub(S, D);
      } else {
        const double2* in = reinterpret_cast<const double2*>(X) + b * a.nring * (int64_t)a.mmax;
        G = in[(int64_t)(side ? a.nring - 1 - k : k) * a.mmax + m];
      }
      const bool edge = (m == 0) || (2 * m == N);
      double c = wscale;
      if (a.cmode == 1 && !edge) c *= 0.5;
      double2 x = zmul(zscale(G, c), zconj(ph[m]));
      if (edge) x.y = 0.0;
      return x;
    };
    auto uval = [&](int j) -> double2 {
      const double2 f1 = spec(j);
      const double2 f2 = zconj(spec(H - j));
      const double2 Ep = zadd(f1, f2);
      double sn, cs;
      sincospi(2.0 * j / N, &sn, &cs);
      const double2 Op = zmul(zsub(f1, f2), make_double2(cs, sn));
      return zconj(make_double2(Ep.x - Op.y, Ep.y + Op.x));
    };
    double* dst = fr + (side ? pr.off_s : pr.off_n);
    if (pr.half == 1) {
      ring_seq_dft64(buf, a, pr, uval);
      for (int j = threadIdx.x; j < H; j += kT64) {
        const double2 r = buf[j];
        dst[2 * j] = r.x;
        dst[2 * j + 1] = -r.y;
      }
    } else {
      double2 Areg[kHalfRegs64];
      ring_seq_dft64(buf, a, pr, [&](int j) { return uval(2 * j); });
#pragma unroll
      for (int i = 0; i < kHalfRegs64; ++i) {
        const int x = threadIdx.x + i * kT64;
        if (x < Q) Areg[i] = buf[x];
      }
      __syncthreads();
      ring_seq_dft64(buf, a, pr, [&](int j) { return uval(2 * j + 1); });
#pragma unroll
      for (int i = 0; i < kHalfRegs64; ++i) {
        const int x = threadIdx.x + i * kT64;
        if (x < Q) {
          double sn, cs;
          sincospi(-2.0 * x / H, &sn, &cs);
          const double2 wb = zmul(make_double2(cs, sn), buf[x]);
          const double2 r0 = zadd(Areg[i], wb), r1 = zsub(Areg[i], wb);
          dst[2 * x] = r0.x;
          dst[2 * x + 1] = -r0.y;
          dst[2 * (x + Q)] = r1.x;
          dst[2 * (x + Q) + 1] = -r1.y;
        }
      }
    }
    __syncthreads();
  }
}

// One ring pair x one map per work-group.  layout 0: folded spectra
// X[m][parity][k][2 b + re/im] (every m < mmax written, zeros from
// K = min(mmax, N/2+1) on); layout 1: ring-major ftm[b][t][m] (complex).
// cmode 1 doubles the orders 0 < m < N/2 (the exact adjoint of the
// synthesis), as in the FP32 kernels.
__global__ __launch_bounds__(kT64) void ring_analysis64_kernel(const double* __restrict__ f,
                                                               double* __restrict__ X, Args64 a) {
  extern __shared__ double2 buf64[];
  const int k = (int)(blockIdx.x / a.nbatch);
  const int64_t b = blockIdx.x % a.nbatch;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const double* fr = f + b * a.npix;
  if (pr.half) {
    analysis_half64(buf64, a, pr, k, b, fr, X);
    return;
  }
  for (int j = threadIdx.x; j < M; j += kT64) {
    double2 z = make_double2(0.0, 0.0);
    if (j < N) {
      z.x = fr[pr.off_n + j];
      z.y = pr.single ? 0.0 : fr[pr.off_s + j];
    }
    buf64[j] = z;
  }
  __syncthreads();
  ring_dft64(buf64, pr, a.chirp + pr.chirp_off, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
  const double wscale = (a.use_weight ? a.w[k] : 1.0) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const double2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int m = threadIdx.x; m < a.mmax; m += kT64) {
    double2 Fn = make_double2(0.0, 0.0), Fs = Fn;
    if (m < K) {
      const double c = (a.cmode == 1 && m != 0 && 2 * m != N) ? 2.0 * wscale : wscale;
      const double2 zm = buf64[m], zc = buf64[m == 0 ? 0 : N - m];
      const double2 xn = make_double2(0.5 * (zm.x + zc.x), 0.5 * (zm.y - zc.y));
      const double2 xs = make_double2(0.5 * (zm.y + zc.y), 0.5 * (zc.x - zm.x));
      Fn = zscale(zmul(xn, ph[m]), c);
      Fs = zscale(zmul(xs, ph[m]), c);
    }
    if (a.layout == 0) {
      const double2 S = pr.single ? Fn : zadd(Fn, Fs);
      const double2 D = pr.single ? Fn : zsub(Fn, Fs);
      double* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
      *reinterpret_cast<double2*>(base) = S;
      *reinterpret_cast<double2*>(base + pstride) = D;
    } else {
      double2* out = reinterpret_cast<double2*>(X) + b * a.nring * (int64_t)a.mmax;
      out[(int64_t)k * a.mmax + m] = Fn;
      if (!pr.single) out[(int64_t)(a.nring - 1 - k) * a.mmax + m] = Fs;
    }
  }
}

// Inverse of the above: per ring, irfft(norm="forward") of the ring's
// spectrum times exp(+i m phi0) (cmode 1: the orders 0 < m < N/2 halved, the
// operator of the exact adjoint of the analysis).
__global__ __launch_bounds__(kT64) void ring_synthesis64_kernel(const double* __restrict__ X,
                                                                double* __restrict__ f, Args64 a) {
  extern __shared__ double2 buf64[];
  const int k = (int)(blockIdx.x / a.nbatch);
  const int64_t b = blockIdx.x % a.nbatch;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const double wscale = (a.use_weight ? a.w[k] : 1.0) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const double2* ph = a.phase + (int64_t)k * a.mmax;
  if (pr.half) {
    synthesis_half64(buf64, a, pr, k, b, X, f + b * a.npix);
    return;
  }
  for (int x = threadIdx.x; x < M; x += kT64) buf64[x] = make_double2(0.0, 0.0);
  __syncthreads();
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int m = threadIdx.x; m < K; m += kT64) {
    double2 Gn, Gs;
    if (a.layout == 0) {
      const double* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
      const double2 S = *reinterpret_cast<const double2*>(base);
      const double2 D = *reinterpret_cast<const double2*>(base + pstride);
      Gn = zadd(S, D);
      Gs = zsub(S, D);
    } else {
      const double2* in = reinterpret_cast<const double2*>(X) + b * a.nring * (int64_t)a.mmax;
      Gn = in[(int64_t)k * a.mmax + m];
      Gs = pr.single ? Gn : in[(int64_t)(a.nring - 1 - k) * a.mmax + m];
    }
    // Hermitian packing of the two rings' spectra (conjugated: the forward
    // DFT of conj(Z) gives conj of the inverse DFT)
    const bool edge = (m == 0) || (2 * m == N);
    double c = wscale;
    if (a.cmode == 1 && !edge) c *= 0.5;
    const double2 phc = zconj(ph[m]);
    double2 xn = zmul(zscale(Gn, c), phc);
    double2 xs = pr.single ? make_double2(0.0, 0.0) : zmul(zscale(Gs, c), phc);
    if (edge) {
      xn.y = 0.0;
      xs.y = 0.0;
    }
    buf64[m] = zconj(make_double2(xn.x - xs.y, xn.y + xs.x));
    if (!edge) buf64[N - m] = zconj(make_double2(xn.x + xs.y, xs.x - xn.y));
  }
  __syncthreads();
  ring_dft64(buf64, pr, a.chirp + pr.chirp_off, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
  double* fr = f + b * a.npix;
  for (int j = threadIdx.x; j < N; j += kT64) {
    const double2 r = buf64[j];
    fr[pr.off_n + j] = r.x;
    if (!pr.single) fr[pr.off_s + j] = -r.y;
  }
}

// ---------------------------------------------------------------------------
// FP64 GEMMs per (m, parity): 32 rows x 64 columns per work-group, 256
// threads with 2 x 4 outputs each, reduction in chunks of 16 through LDS.
// ---------------------------------------------------------------------------

// Column signs of the vector transform: columns come in fours per field,
// (component, re / im); the D0 half of the reduction takes column c with sign
// (+, +, -, -), the D1 half column c ^ 3 with sign (+, -, +, -) forward and
// (-, +, -, +) inverse (hpx_sht.py:351-372 and 413-432 written out per real
// column, as in the FP32 vector GEMMs).
__device__ __forceinline__ double vsign0(int64_t c) { return (c & 2) ? -1.0 : 1.0; }
__device__ __forceinline__ double vsign1(int64_t c, bool forward) { return ((c & 1) != 0) == forward ? -1.0 : 1.0; }

// Forward: aT[row(m, p, i)][col] = sum_{k < npair} P[m][p][i][k] X[m][p][k][col]
// for rows i < nl[m][p]; items (m, p << 16 | row tile), blockIdx.y = column tile.
// VEC (vector transform): the sum over k runs twice, the D0 table (at P)
// against the opposite parity's spectra, then the D1 table (at P + vstride)
// against the same parity's, with the column map and signs above.
template <bool VEC>
__global__ __launch_bounds__(kGemmThreads) void gemm_fwd64_kernel(const double* __restrict__ X,
                                                                  double* __restrict__ aT,
                                                                  const double* __restrict__ P,
                                                                  const int64_t* __restrict__ p_off,
                                                                  const int* __restrict__ nl,
                                                                  const int2* __restrict__ items, int kpad,
                                                                  int npair, int64_t ncol, int64_t vstride) {
  __shared__ double As[32][17];
  __shared__ double Bs[16][65];
  const int2 it = items[blockIdx.x];
  const int m = it.x, p = it.y >> 16, rt = it.y & 0xffff;
  const int64_t c0 = (int64_t)blockIdx.y * 64;
  const int nlp = nl[2 * m + p];
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  for (int half = 0; half < (VEC ? 2 : 1); ++half) {
  const double* Pm = P + half * vstride + p_off[m] + (p ? (int64_t)pad0 * kpad : 0);
  const int pb = VEC && half == 0 ? 1 - p : p;  // spectra parity block
  const double* Xm = X + ((int64_t)m * 2 + pb) * kpad * ncol;
  for (int k0 = 0; k0 < npair; k0 += 16) {
    for (int e = tid; e < 32 * 16; e += kGemmThreads) {
      const int r = e >> 4, kk = e & 15, i = rt * 32 + r;
      As[r][kk] = (i < nlp && k0 + kk < npair) ? Pm[(int64_t)i * kpad + k0 + kk] : 0.0;
    }
    for (int e = tid; e < 16 * 64; e += kGemmThreads) {
      const int kk = e >> 6, c = e & 63;
      const int64_t cc = VEC && half == 1 ? (c0 + c) ^ 3 : c0 + c;
      const double sg = VEC ? (half == 0 ? vsign0(c0 + c) : vsign1(c0 + c, true)) : 1.0;
      Bs[kk][c] = (k0 + kk < npair && c0 + c < ncol) ? sg * Xm[(int64_t)(k0 + kk) * ncol + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      const double a0 = As[2 * ty][kk], a1 = As[2 * ty + 1][kk];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double bv = Bs[kk][tx + 16 * j];
        acc[0][j] = fma(a0, bv, acc[0][j]);
        acc[1][j] = fma(a1, bv, acc[1][j]);
      }
    }
    __syncthreads();
  }
  }
  const int64_t row0 = p_off[m] / kpad + (p ? pad0 : 0);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int i = rt * 32 + 2 * ty + q;
    if (i >= nlp) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t c = c0 + tx + 16 * j;
      if (c < ncol) aT[(row0 + i) * ncol + c] = acc[q][j];
    }
  }
}

// Inverse: X[m][p][k][col] = sum_{i < nl[m][p]} P[m][p][i][k] aT[row(m, p, i)][col]
// for ring pairs k < npair; blockIdx.x = (m * 2 + p) * nkt + ring-pair tile,
// blockIdx.y = column tile.  VEC: the sum runs over the D0 table of the
// opposite parity against that parity's coefficients, then over the D1 table
// (at P + vstride) of the same parity, with the column map and signs above.
template <bool VEC>
__global__ __launch_bounds__(kGemmThreads) void gemm_inv64_kernel(const double* __restrict__ aT,
                                                                  double* __restrict__ X,
                                                                  const double* __restrict__ P,
                                                                  const int64_t* __restrict__ p_off,
                                                                  const int* __restrict__ nl, int nkt, int kpad,
                                                                  int npair, int64_t ncol, int64_t vstride) {
  __shared__ double As[16][33];
  __shared__ double Bs[16][65];
  const int mp = blockIdx.x / nkt, kt = blockIdx.x % nkt;
  const int m = mp >> 1, p = mp & 1;
  const int64_t c0 = (int64_t)blockIdx.y * 64;
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  const int tid = threadIdx.x, ty = tid >> 4, tx = tid & 15;
  const int k0 = kt * 32;
  double acc[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
  for (int half = 0; half < (VEC ? 2 : 1); ++half) {
  const int tp = VEC && half == 0 ? 1 - p : p;  // table and coefficient parity
  const int nlp = nl[2 * m + tp];
  const double* Pm = P + half * vstride + p_off[m] + (tp ? (int64_t)pad0 * kpad : 0);
  const int64_t row0 = p_off[m] / kpad + (tp ? pad0 : 0);
  for (int i0 = 0; i0 < nlp; i0 += 16) {
    for (int e = tid; e < 16 * 32; e += kGemmThreads) {
      const int ii = e >> 5, r = e & 31, k = k0 + r;
      As[ii][r] = (i0 + ii < nlp && k < npair) ? Pm[(int64_t)(i0 + ii) * kpad + k] : 0.0;
    }
    for (int e = tid; e < 16 * 64; e += kGemmThreads) {
      const int ii = e >> 6, c = e & 63;
      const int64_t cc = VEC && half == 1 ? (c0 + c) ^ 3 : c0 + c;
      const double sg = VEC ? (half == 0 ? vsign0(c0 + c) : vsign1(c0 + c, false)) : 1.0;
      Bs[ii][c] = (i0 + ii < nlp && c0 + c < ncol) ? sg * aT[(row0 + i0 + ii) * ncol + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ii = 0; ii < 16; ++ii) {
      const double a0 = As[ii][2 * ty], a1 = As[ii][2 * ty + 1];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double bv = Bs[ii][tx + 16 * j];
        acc[0][j] = fma(a0, bv, acc[0][j]);
        acc[1][j] = fma(a1, bv, acc[1][j]);
      }
    }
    __syncthreads();
  }
  }
  const int64_t pstride = (int64_t)kpad * ncol;
  double* Xm = X + ((int64_t)m * 2 + p) * pstride;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int k = k0 + 2 * ty + q;
    if (k >= npair) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t c = c0 + tx + 16 * j;
      if (c < ncol) Xm[(int64_t)k * ncol + c] = acc[q][j];
    }
  }
}

__device__ __forceinline__ int64_t at_row64(const int64_t* p_off, const int* nl, int kpad, int m, int l) {
  const int d = l - m;
  const int pad0 = (nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
  return p_off[m] / kpad + ((d & 1) ? pad0 : 0) + (d >> 1);
}

// a[b][l][m] (complex128) from aT; zero for m > l.  One thread per (b, l, m).
// lnorm: scale by 1 / (l (l + 1)) (1 at l = 0), the vector transform's
// normalisation (hpx_sht.py:323-328).
__global__ __launch_bounds__(256) void coeff_to_user64_kernel(const double* __restrict__ aT,
                                                              double* __restrict__ alm,
                                                              const int64_t* __restrict__ p_off,
                                                              const int* __restrict__ nl, int kpad, int lmax,
                                                              int mmax, int64_t ncol, int64_t nbatch, int lnorm) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = nbatch * lmax * (int64_t)mmax;
  if (e >= total) return;
  const int m = (int)(e % mmax);
  const int l = (int)((e / mmax) % lmax);
  const int64_t b = e / ((int64_t)mmax * lmax);
  double2 v = make_double2(0.0, 0.0);
  if (m <= l) {
    const double* src = aT + at_row64(p_off, nl, kpad, m, l) * ncol + 2 * b;
    const double sc = (lnorm && l > 0) ? 1.0 / ((double)l * (l + 1)) : 1.0;
    v = make_double2(sc * src[0], sc * src[1]);
  }
  reinterpret_cast<double2*>(alm)[e] = v;
}

// aT from a[b][l][m] (rows m <= l only).
__global__ __launch_bounds__(256) void user_to_coeff64_kernel(const double* __restrict__ alm,
                                                              double* __restrict__ aT,
                                                              const int64_t* __restrict__ p_off,
                                                              const int* __restrict__ nl, int kpad, int lmax,
                                                              int mmax, int64_t ncol, int64_t nbatch) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t total = nbatch * lmax * (int64_t)mmax;
  if (e >= total) return;
  const int m = (int)(e % mmax);
  const int l = (int)((e / mmax) % lmax);
  const int64_t b = e / ((int64_t)mmax * lmax);
  if (m > l) return;
  const double2 v = reinterpret_cast<const double2*>(alm)[e];
  double* dst = aT + at_row64(p_off, nl, kpad, m, l) * ncol + 2 * b;
  dst[0] = v.x;
  dst[1] = v.y;
}

template <typename T>
int upload64(T** dst, const std::vector<T>& src) {
  if (src.empty()) return HPX_OK;
  HPX_HIP(hipMalloc(reinterpret_cast<void**>(dst), src.size() * sizeof(T)));
  HPX_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return HPX_OK;
}

inline int64_t x_doubles(const ShtPlan& p, int64_t nbatch) {
  return (int64_t)p.mmax * 2 * p.kpad * col_pad(nbatch);
}
inline int64_t workspace_doubles(const ShtPlan& p, int64_t nbatch) {
  return x_doubles(p, nbatch) + p.at_rows * col_pad(nbatch);
}

Args64 make_args64(const ShtPlan& p, int64_t nbatch, int use_weight, double scale, int cmode, int layout) {
  Args64 a;
  a.pairs = p.d_pairs64;
  a.phase = p.d_phase64;
  a.chirp = p.d_chirp64;
  a.bhat = p.d_bhat64;
  a.tw = p.d_tw64;
  a.w = p.d_w64;
  a.log2_tab = p.log2_max_M64;
  a.npix = p.npix;
  a.nbatch = nbatch;
  a.ncol = col_pad(nbatch);
  a.mmax = p.mmax;
  a.kpad = p.kpad;
  a.nring = p.nring;
  a.use_weight = use_weight;
  a.cmode = cmode;
  a.layout = layout;
  a.scale = scale;
  return a;
}

int launch_fft64(const ShtPlan& p, bool analysis, const double* in, double* out, int64_t nbatch, int use_weight,
                 double scale, int cmode, int layout, hipStream_t s) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t items = (int64_t)p.npair * nbatch;
  if (items > 0x7fffffff) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const size_t lds = (size_t)p.max_M64 * sizeof(double2);
  static const bool attr = [] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ring_analysis64_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ring_synthesis64_kernel),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    return true;
  }();
  (void)attr;
  const Args64 a = make_args64(p, nbatch, use_weight, scale, cmode, layout);
  if (analysis)
    ring_analysis64_kernel<<<(unsigned)items, kT64, lds, s>>>(in, out, a);
  else
    ring_synthesis64_kernel<<<(unsigned)items, kT64, lds, s>>>(in, out, a);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_legendre_fwd64(const ShtPlan& p, double* ws, double* alm, int64_t nbatch, hipStream_t s) {
  const int64_t ncol = col_pad(nbatch);
  double* aT = ws + x_doubles(p, nbatch);
  const int64_t nct = (ncol + 63) / 64;
  if (p.n_items64 > 0) {
    if (nct > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
    const dim3 g((unsigned)p.n_items64, (unsigned)nct);
    if (p.vector)
      gemm_fwd64_kernel<true><<<g, kGemmThreads, 0, s>>>(ws, aT, p.d_p64, p.d_p_off, p.d_nl, p.d_items64, p.kpad,
                                                          p.npair, ncol, p.vstride);
    else
      gemm_fwd64_kernel<false><<<g, kGemmThreads, 0, s>>>(ws, aT, p.d_p64, p.d_p_off, p.d_nl, p.d_items64, p.kpad,
                                                           p.npair, ncol, 0);
    HPX_LAUNCH_CHECK();
  }
  const int64_t total = nbatch * p.lmax * (int64_t)p.mmax;
  if ((total + 255) / 256 > 0x7fffffff) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  coeff_to_user64_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(aT, alm, p.d_p_off, p.d_nl, p.kpad, p.lmax,
                                                                        p.mmax, ncol, nbatch, p.vector);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_legendre_inv64(const ShtPlan& p, const double* alm, double* ws, int64_t nbatch, hipStream_t s) {
  const int64_t ncol = col_pad(nbatch);
  double* aT = ws + x_doubles(p, nbatch);
  const int64_t total = nbatch * p.lmax * (int64_t)p.mmax;
  if ((total + 255) / 256 > 0x7fffffff) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  user_to_coeff64_kernel<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(alm, aT, p.d_p_off, p.d_nl, p.kpad, p.lmax,
                                                                        p.mmax, ncol, nbatch);
  HPX_LAUNCH_CHECK();
  const int nkt = (p.npair + 31) / 32;
  const int64_t nx = (int64_t)p.mmax * 2 * nkt;
  const int64_t nct = (ncol + 63) / 64;
  if (nx > 0x7fffffff || nct > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const dim3 g((unsigned)nx, (unsigned)nct);
  if (p.vector)
    gemm_inv64_kernel<true><<<g, kGemmThreads, 0, s>>>(aT, ws, p.d_p64, p.d_p_off, p.d_nl, nkt, p.kpad, p.npair, ncol,
                                                       p.vstride);
  else
    gemm_inv64_kernel<false><<<g, kGemmThreads, 0, s>>>(aT, ws, p.d_p64, p.d_p_off, p.d_nl, nkt, p.kpad, p.npair,
                                                        ncol, 0);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

// nbatch: scalar maps (2 per field of a vector call)
int check64(const ShtPlan* p, int64_t nbatch, int64_t ws_bytes, bool need_table, bool vector_call = false) {
  if (p == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (p->vector && !vector_call) HPX_FAIL(HPX_ERR_ARG, "plan holds vector-transform tables (use hpx_vsht_*_f64)");
  if (!p->vector && vector_call) HPX_FAIL(HPX_ERR_ARG, "plan holds the scalar table (use a vector plan)");
  if (!p->f64) HPX_FAIL(HPX_ERR_ARG, "plan has no FP64 tables: call hpx_sht_plan_enable_f64 first");
  if (need_table && p->lmax < 1) HPX_FAIL(HPX_ERR_ARG, "plan was built with lmax = 0 (ring FFTs only)");
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch > max_maps(*p) / 2)
    HPX_FAIL(HPX_ERR_UNSUPPORTED, "%lld maps in one FP64 call exceed this plan's limit of %lld; split the batch",
             (long long)nbatch, (long long)(max_maps(*p) / 2));
  if (ws_bytes >= 0) {
    const int64_t need = workspace_doubles(*p, nbatch) * (int64_t)sizeof(double);
    if (ws_bytes < need)
      HPX_FAIL(HPX_ERR_ARG, "workspace of %lld bytes is smaller than the %lld bytes needed", (long long)ws_bytes,
               (long long)need);
  }
  return HPX_OK;
}

}  // namespace
}  // namespace hpxmi

using namespace hpxmi;

namespace {
ShtPlan* P64(const hpx_sht_plan* h) { return reinterpret_cast<ShtPlan*>(const_cast<hpx_sht_plan*>(h)); }
int ceil_log2(int64_t n) {
  int r = 0;
  while ((int64_t(1) << r) < n) ++r;
  return r;
}
}  // namespace

extern "C" int hpx_sht_plan_enable_f64(hpx_sht_plan* plan, hipStream_t stream) {
  ShtPlan* p = P64(plan);
  if (p == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  std::lock_guard<std::mutex> g(p->tune_mu);
  if (p->f64) return HPX_OK;
  // the FP64 kernels' own FFT length per pair (at most kMaxM64 points): the
  // pair whole, or ring by ring at half or quarter length as in FP32
  std::vector<RingPair> pairs = p->pairs;
  int max64 = 1;
  for (int k = 0; k < p->npair; ++k) {
    RingPair& r = pairs[k];
    const int64_t N = r.nphi;
    auto fft_len = [](int64_t n) { return (n & (n - 1)) == 0 ? n : int64_t(1) << ceil_log2(2 * n - 1); };
    r.half = 0;
    int64_t M = fft_len(N);
    if (M > kMaxM64 && N % 2 == 0 && fft_len(N / 2) <= kMaxM64) {
      r.half = 1;
      M = fft_len(N / 2);
    } else if (M > kMaxM64 && N % 4 == 0 && ((N / 4) & (N / 4 - 1)) != 0 && fft_len(N / 4) <= kMaxM64) {
      r.half = 2;
      M = fft_len(N / 4);
    }
    if (M > kMaxM64)
      HPX_FAIL(HPX_ERR_UNSUPPORTED, "FP64 ring FFTs are limited to %d points (a ring of %lld pixels needs %lld)",
               kMaxM64, (long long)N, (long long)M);
    r.M = (int)M;
    r.log2M = ceil_log2(M);
    r.nseq = (int)(N >> r.half);
    max64 = std::max(max64, r.M);
  }
  const int mmax = p->mmax;
  // scalar plans: P; vector plans: D0 and D1 (D1 at vstride)
  const int64_t table = std::max<int64_t>((p->vector ? 2 : 1) * p->p_off[mmax], 1);
  {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (uint64_t)(table * sizeof(double)) > (uint64_t)free_b)
      HPX_FAIL(HPX_ERR_UNSUPPORTED, "the FP64 Legendre table needs %.2f GB, %.2f GB free on the device",
               table * sizeof(double) / 1e9, free_b / 1e9);
  }
  // Bluestein tables, twiddles, phases and weights in FP64 (the FP32 plan's
  // formulas, kept in FP64): chirp exp(-i pi (j^2 mod 2N) / N), kernel
  // spectrum FFT_M(conj chirp) / M
  std::vector<double2> chirp, bhat;
  for (int k = 0; k < p->npair; ++k) {
    RingPair& r = pairs[k];
    r.chirp_off = 0;
    r.bhat_off = 0;
    if (r.M == r.nseq) continue;
    const int N = r.nseq;
    r.chirp_off = (int)chirp.size();
    r.bhat_off = (int)bhat.size();
    std::vector<cd> w(N), b(r.M, cd(0.0, 0.0));
    for (int j = 0; j < N; ++j) {
      const int64_t q = ((int64_t)j * j) % (2 * (int64_t)N);
      const double ang = -M_PI * (double)q / (double)N;
      w[j] = cd(std::cos(ang), std::sin(ang));
      chirp.push_back(make_double2(w[j].real(), w[j].imag()));
    }
    b[0] = std::conj(w[0]);
    for (int j = 1; j < N; ++j) {
      b[j] = std::conj(w[j]);
      b[r.M - j] = std::conj(w[j]);
    }
    host_fft(b);
    for (int x = 0; x < r.M; ++x) bhat.push_back(make_double2(b[x].real() / r.M, b[x].imag() / r.M));
  }
  if (chirp.empty()) {
    chirp.push_back(make_double2(1.0, 0.0));
    bhat.push_back(make_double2(1.0, 0.0));
  }
  std::vector<double2> tw(max64), phase((size_t)p->npair * mmax);
  for (int x = 0; x < max64; ++x) {
    const double ang = -2.0 * M_PI * (double)x / (double)max64;
    tw[x] = make_double2(std::cos(ang), std::sin(ang));
  }
  for (int k = 0; k < p->npair; ++k)
    for (int m = 0; m < mmax; ++m) {
      const double ang = std::fmod((double)m * p->h_phi0[k], 2.0 * M_PI);
      phase[(size_t)k * mmax + m] = make_double2(std::cos(ang), -std::sin(ang));
    }
  std::vector<int2> items;
  for (int m = 0; m < mmax; ++m)
    for (int par = 0; par < 2; ++par)
      for (int rt = 0; rt * 32 < p->nl[2 * m + par]; ++rt) items.push_back(make_int2(m, (par << 16) | rt));
  // the FP64 kernels read their own copy of the ring pairs, whose chirp and
  // kernel-spectrum offsets index the FP64 tables
  RingPair* d_pairs64 = nullptr;
  int rc = HPX_OK;
  double2 *d_chirp = nullptr, *d_bhat = nullptr, *d_tw = nullptr, *d_phase = nullptr;
  double* d_w = nullptr;
  int2* d_items = nullptr;
  double* d_p = nullptr;
  double* d_ct = nullptr;
  auto fail = [&](int code) {
    hipFree(d_pairs64);
    hipFree(d_chirp);
    hipFree(d_bhat);
    hipFree(d_tw);
    hipFree(d_phase);
    hipFree(d_w);
    hipFree(d_items);
    hipFree(d_p);
    hipFree(d_ct);
    return code;
  };
  std::vector<double> ct(p->kpad, 0.0);
  for (int k = 0; k < p->npair; ++k) ct[k] = p->h_ct[k];
  if ((rc = upload64(&d_pairs64, pairs)) != HPX_OK || (rc = upload64(&d_chirp, chirp)) != HPX_OK ||
      (rc = upload64(&d_bhat, bhat)) != HPX_OK || (rc = upload64(&d_tw, tw)) != HPX_OK ||
      (rc = upload64(&d_phase, phase)) != HPX_OK || (rc = upload64(&d_w, p->h_w)) != HPX_OK ||
      (rc = upload64(&d_items, items)) != HPX_OK || (rc = upload64(&d_ct, ct)) != HPX_OK)
    return fail(rc);
  if (hipMalloc(reinterpret_cast<void**>(&d_p), table * sizeof(double)) != hipSuccess) {
    set_error("FP64 Legendre table allocation failed");
    return fail(HPX_ERR_HIP);
  }
  if (mmax > 0 && p->lmax > 0) {
    if (p->vector) {
      if (launch_vector_table64(*p, d_p, d_ct, stream) != HPX_OK) return fail(HPX_ERR_HIP);
    } else {
      legendre_table64_kernel<<<dim3((unsigned)((p->kpad + 255) / 256), (unsigned)mmax), 256, 0, stream>>>(
          d_p, p->d_p_off, p->d_nl, d_ct, p->npair, p->kpad, p->lmax);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(stream) != hipSuccess) {
      set_error("FP64 Legendre table generation failed");
      return fail(HPX_ERR_HIP);
    }
  }
  hipFree(d_ct);
  p->d_chirp64 = d_chirp;
  p->d_bhat64 = d_bhat;
  p->d_tw64 = d_tw;
  p->d_phase64 = d_phase;
  p->d_w64 = d_w;
  p->d_items64 = d_items;
  p->n_items64 = (int)items.size();
  p->d_p64 = d_p;
  p->d_pairs64 = d_pairs64;
  p->max_M64 = max64;
  p->log2_max_M64 = ceil_log2(max64);
  p->table_bytes += table * (int64_t)sizeof(double);
  p->f64 = 1;
  return HPX_OK;
}

extern "C" int64_t hpx_sht_workspace_bytes_f64(const hpx_sht_plan* plan, int64_t nbatch) {
  if (plan == nullptr || nbatch < 0) return -1;
  return workspace_doubles(*P64(plan), nbatch) * (int64_t)sizeof(double);
}

extern "C" int hpx_sht_forward_f64(const hpx_sht_plan* plan, const double* f, double* alm, int64_t nbatch,
                                   void* workspace, int64_t workspace_bytes, hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  double* ws = static_cast<double*>(workspace);
  if ((rc = launch_fft64(*pl, true, f, ws, nbatch, 1, 1.0, 0, 0, stream)) != HPX_OK) return rc;
  return launch_legendre_fwd64(*pl, ws, alm, nbatch, stream);
}

extern "C" int hpx_sht_inverse_f64(const hpx_sht_plan* plan, const double* alm, double* f, int64_t nbatch,
                                   void* workspace, int64_t workspace_bytes, hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  double* ws = static_cast<double*>(workspace);
  if ((rc = launch_legendre_inv64(*pl, alm, ws, nbatch, stream)) != HPX_OK) return rc;
  return launch_fft64(*pl, false, ws, f, nbatch, 0, 1.0, 0, 0, stream);
}

extern "C" int hpx_sht_forward_adjoint_f64(const hpx_sht_plan* plan, int mode, const double* galm, double* gf,
                                           int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                           hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  if (mode != HPX_ADJ_EXACT && mode != HPX_ADJ_COMPAT) HPX_FAIL(HPX_ERR_ARG, "unknown adjoint mode %d", mode);
  double* ws = static_cast<double*>(workspace);
  if ((rc = launch_legendre_inv64(*pl, galm, ws, nbatch, stream)) != HPX_OK) return rc;
  // exact: w_t * sum_{m<K} Re(H e^{i m phi});  compat: (w_t / 2) * irfft(H)
  if (mode == HPX_ADJ_EXACT) return launch_fft64(*pl, false, ws, gf, nbatch, 1, 1.0, 1, 0, stream);
  return launch_fft64(*pl, false, ws, gf, nbatch, 1, 0.5, 0, 0, stream);
}

extern "C" int hpx_sht_inverse_adjoint_f64(const hpx_sht_plan* plan, int mode, const double* gf, double* galm,
                                           int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                           hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  if (mode != HPX_ADJ_EXACT && mode != HPX_ADJ_COMPAT) HPX_FAIL(HPX_ERR_ARG, "unknown adjoint mode %d", mode);
  double* ws = static_cast<double*>(workspace);
  // exact: c_m * rfft(g);  compat: 2 * rfft(g)
  if (mode == HPX_ADJ_EXACT) rc = launch_fft64(*pl, true, gf, ws, nbatch, 0, 1.0, 1, 0, stream);
  else rc = launch_fft64(*pl, true, gf, ws, nbatch, 0, 2.0, 0, 0, stream);
  if (rc != HPX_OK) return rc;
  return launch_legendre_fwd64(*pl, ws, galm, nbatch, stream);
}

extern "C" int hpx_ring_rfft_f64(const hpx_sht_plan* plan, const double* f, double* ftm, int64_t nbatch,
                                 hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, -1, false);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_fft64(*pl, true, f, ftm, nbatch, 0, 1.0, 0, 1, stream);
}

extern "C" int hpx_ring_irfft_f64(const hpx_sht_plan* plan, const double* ftm, double* f, int64_t nbatch,
                                  hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  int rc = check64(pl, nbatch, -1, false);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_fft64(*pl, false, ftm, f, nbatch, 0, 1.0, 0, 1, stream);
}

extern "C" int64_t hpx_vsht_workspace_bytes_f64(const hpx_sht_plan* plan, int64_t nbatch) {
  if (plan == nullptr || nbatch < 0 || nbatch > (INT64_MAX >> 3)) return -1;
  return workspace_doubles(*P64(plan), 2 * nbatch) * (int64_t)sizeof(double);
}

extern "C" int hpx_vsht_forward_f64(const hpx_sht_plan* plan, const double* f, double* alm, int64_t nbatch,
                                    void* workspace, int64_t workspace_bytes, hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  if (nbatch < 0 || nbatch > (INT64_MAX >> 3)) HPX_FAIL(HPX_ERR_ARG, "bad nbatch");
  int rc = check64(pl, 2 * nbatch, workspace_bytes, true, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  // the two components of a field are consecutive maps: 2 nbatch ring analyses
  double* ws = static_cast<double*>(workspace);
  if ((rc = launch_fft64(*pl, true, f, ws, 2 * nbatch, 1, 1.0, 0, 0, stream)) != HPX_OK) return rc;
  return launch_legendre_fwd64(*pl, ws, alm, 2 * nbatch, stream);
}

extern "C" int hpx_vsht_inverse_f64(const hpx_sht_plan* plan, const double* alm, double* f, int64_t nbatch,
                                    void* workspace, int64_t workspace_bytes, hipStream_t stream) {
  ShtPlan* pl = P64(plan);
  if (nbatch < 0 || nbatch > (INT64_MAX >> 3)) HPX_FAIL(HPX_ERR_ARG, "bad nbatch");
  int rc = check64(pl, 2 * nbatch, workspace_bytes, true, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  double* ws = static_cast<double*>(workspace);
  if ((rc = launch_legendre_inv64(*pl, alm, ws, 2 * nbatch, stream)) != HPX_OK) return rc;
  return launch_fft64(*pl, false, ws, f, 2 * nbatch, 0, 1.0, 0, 0, stream);
}
