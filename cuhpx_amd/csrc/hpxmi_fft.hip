// Per-ring-pair FFT kernels of the SHT (gfx950).
//
// Analysis (forward SHT's first stage, and the adjoint of the inverse SHT)
// replaces cuhpx_fft.healpix_rfft_batch (src/harmonic_transform/hpx_fft.cpp:21-84):
//   F_t[m] = e^{-i m phi0_t} sum_j f[off_t + j] e^{-2 pi i m j / N_t}
//   for m < K_t = min(mmax, N_t/2 + 1) and 0 above (hpx_fft_cuda.cu:1107-1178).
// Synthesis (inverse SHT's last stage, and the adjoint of the forward SHT)
// replaces cuhpx_fft.healpix_irfft_batch (hpx_fft.cpp:86-149):
//   f[off_t + j] = sum_{m < K_t} c_m Re(G_t[m] e^{i m (phi0_t + 2 pi j / N_t)})
//   with irfft(norm="forward") semantics (hpx_fft_cuda.cu:1213-1332).
//
// MI355X design.  One work-group owns one mirror pair of rings (north t and
// south nring-1-t have the same length, phase and weight) and 8 fields.  The
// two real rings of a pair are packed into one complex sequence z = f_t + i
// f_t', so one complex DFT serves both.  The DFT runs entirely in LDS: a
// radix-8/4/2 Stockham FFT (each pass: registers <- LDS, barrier, twiddle +
// butterfly, LDS <- registers, barrier).  Rings whose length is not a power
// of two use Bluestein's chirp-z algorithm with a padded power-of-two length
// M >= 2N-1 and FP64-built chirp / kernel-spectrum tables.  Chirp multiply,
// phase shift, quadrature weight, the split into the two rings and the
// equatorial folding S = F_t + F_t', D = F_t - F_t' (the Legendre step's
// operands for l+m even / odd) are fused into the load and store loops, so
// the only HBM traffic is the map and the folded spectra.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

namespace hpxmi {
namespace {

constexpr int kThreads = 512;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by -i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

template <int R>
__device__ __forceinline__ void dft(float2* v);

template <>
__device__ __forceinline__ void dft<2>(float2* v) {
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void dft<4>(float2* v) {
  float2 t0 = cadd(v[0], v[2]);
  float2 t1 = csub(v[0], v[2]);
  float2 t2 = cadd(v[1], v[3]);
  float2 t3 = cmul_mi(csub(v[1], v[3]));
  v[0] = cadd(t0, t2);
  v[1] = cadd(t1, t3);
  v[2] = csub(t0, t2);
  v[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void dft<8>(float2* v) {
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft<4>(e);
  dft<4>(o);
  const float h = 0.70710678118654752f;
  // o[q] *= exp(-2 pi i q / 8)
  o[1] = make_float2((o[1].x + o[1].y) * h, (o[1].y - o[1].x) * h);
  o[2] = cmul_mi(o[2]);
  o[3] = make_float2((o[3].y - o[3].x) * h, -(o[3].x + o[3].y) * h);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = cadd(e[q], o[q]);
    v[q + 4] = csub(e[q], o[q]);
  }
}

// One radix-R Stockham pass over `total` points = G sequences of length M
// (M = 2^log2M) stored back to back in buf.  Ns is the length of the
// sub-transforms already combined.  Twiddles come from the table
// tw[x] = exp(-2 pi i x / 2^log2_tab).
template <int R>
__device__ __forceinline__ void stockham_pass(float2* buf, int log2M, int total, int Ns,
                                              int log2Ns, const float2* __restrict__ tw,
                                              int log2_tab) {
  constexpr int kMaxB = kFftLds / R / kThreads;
  constexpr int kLogR = R == 8 ? 3 : (R == 4 ? 2 : 1);
  const int nb = total / R;
  const int log2mr = log2M - kLogR;
  const int mr = 1 << log2mr;
  const int M = 1 << log2M;
  const int tw_shift = log2_tab - log2Ns - kLogR;
  float2 v[kMaxB][R];
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kThreads;
    if (b < nb) {
      const int g = b >> log2mr;
      const int j = b & (mr - 1);
      const float2* src = buf + g * M + j;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = src[r * mr];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kThreads;
    if (b < nb) {
      const int g = b >> log2mr;
      const int j = b & (mr - 1);
      const int jn = j & (Ns - 1);
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[(jn * r) << tw_shift]);
      }
      dft<R>(v[i]);
      float2* dst = buf + g * M + (j - jn) * R + jn;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

// Forward DFT (exp(-2 pi i ...)) of G sequences of length 2^log2M in LDS,
// natural order in and out.  Ends with a barrier.
__device__ void lds_fft(float2* buf, int log2M, int total, const float2* __restrict__ tw,
                        int log2_tab) {
  int Ns = 1, log2Ns = 0, rem = log2M;
  while (rem > 0) {
    if (rem >= 3) {
      stockham_pass<8>(buf, log2M, total, Ns, log2Ns, tw, log2_tab);
      Ns <<= 3;
      log2Ns += 3;
      rem -= 3;
    } else if (rem == 2) {
      stockham_pass<4>(buf, log2M, total, Ns, log2Ns, tw, log2_tab);
      Ns <<= 2;
      log2Ns += 2;
      rem -= 2;
    } else {
      stockham_pass<2>(buf, log2M, total, Ns, log2Ns, tw, log2_tab);
      Ns <<= 1;
      log2Ns += 1;
      rem -= 1;
    }
  }
}

// The DFT of length N of each sequence in buf, entered as z_j (pow2, M == N)
// or z_j * chirp_j zero-padded to M (Bluestein).  On return, for j < N the
// DFT value is dft_value(buf, ...).  Bluestein: X[m] = chirp_m conj(c_m),
// c = FFT(conj(FFT(a) * bhat)) with bhat = FFT(conj chirp)/M.
__device__ void ring_dft(float2* buf, const RingPair& pr, int total, const float2* __restrict__ bhat,
                         const float2* __restrict__ tw, int log2_tab) {
  lds_fft(buf, pr.log2M, total, tw, log2_tab);
  if (pr.M != pr.nphi) {
    const int M = pr.M;
    for (int idx = threadIdx.x; idx < total; idx += kThreads) {
      const int x = idx & (M - 1);
      buf[idx] = cconj(cmul(buf[idx], bhat[x]));
    }
    __syncthreads();
    lds_fft(buf, pr.log2M, total, tw, log2_tab);
  }
}

__device__ __forceinline__ float2 dft_value(const float2* buf, const RingPair& pr, int g, int x,
                                            const float2* __restrict__ chirp) {
  const float2 v = buf[g * pr.M + x];
  if (pr.M == pr.nphi) return v;
  return cmul(chirp[x], cconj(v));
}

__global__ __launch_bounds__(kThreads) void ring_analysis_kernel(
    const float* __restrict__ f, float* __restrict__ X, const RingPair* __restrict__ pairs,
    const float2* __restrict__ phase, const float2* __restrict__ chirp_all,
    const float2* __restrict__ bhat_all, const float2* __restrict__ tw, int log2_tab,
    int64_t npix, int64_t nbatch, int mmax, int kpad, int64_t ncol, int use_weight,
    float scale, int cmode, int layout, int nring) {
  __shared__ float2 buf[kFftLds];
  const int k = blockIdx.x;
  const RingPair pr = pairs[k];
  const int N = pr.nphi;
  const bool blue = pr.M != N;
  const float2* chirp = chirp_all + pr.chirp_off;
  const float2* bhat = bhat_all + pr.bhat_off;
  int log2G = 13 - pr.log2M;
  if (log2G > 3) log2G = 3;
  const int G = 1 << log2G;
  const int64_t b_begin = (int64_t)blockIdx.y * kFieldsPerFftBlock;
  int64_t b_end = b_begin + kFieldsPerFftBlock;
  if (b_end > nbatch) b_end = nbatch;
  const float wscale = (use_weight ? pr.weight : 1.0f) * scale;
  const int K = min(mmax, N / 2 + 1);
  const int64_t pstride = (int64_t)kpad * ncol;
  const int64_t mstride = 2 * pstride;
  const int total = G << pr.log2M;

  for (int64_t b0 = b_begin; b0 < b_end; b0 += G) {
    const int gcount = (int)min((int64_t)G, b_end - b0);
    for (int idx = threadIdx.x; idx < total; idx += kThreads) {
      const int g = idx >> pr.log2M;
      const int j = idx & (pr.M - 1);
      float2 z = make_float2(0.f, 0.f);
      if (g < gcount && j < N) {
        const float* fr = f + (b0 + g) * npix;
        z.x = fr[pr.off_n + j];
        z.y = pr.single ? 0.f : fr[pr.off_s + j];
        if (blue) z = cmul(z, chirp[j]);
      }
      buf[idx] = z;
    }
    __syncthreads();
    ring_dft(buf, pr, total, bhat, tw, log2_tab);

    for (int e = threadIdx.x; e < (mmax << log2G); e += kThreads) {
      // folded layout: fields fastest (consecutive fields are adjacent
      // columns); ring-major layout: m fastest (contiguous along m)
      const int m = layout == 0 ? e >> log2G : e % mmax;
      const int g = layout == 0 ? e & (G - 1) : e / mmax;
      if (g >= gcount) continue;
      float2 S = make_float2(0.f, 0.f), D = make_float2(0.f, 0.f);
      float2 Fn = S, Fs = S;
      if (m < K) {
        const float2 zm = dft_value(buf, pr, g, m, chirp);
        const float2 zc = dft_value(buf, pr, g, m == 0 ? 0 : N - m, chirp);
        // DFTs of the two real rings: Xn = (Z[m] + conj Z[N-m]) / 2,
        // Xs = (Z[m] - conj Z[N-m]) / (2i).
        float2 xn = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y - zc.y));
        float2 xs = make_float2(0.5f * (zm.y + zc.y), 0.5f * (zc.x - zm.x));
        const float2 ph = phase[(int64_t)k * mmax + m];
        xn = cmul(xn, ph);
        xs = cmul(xs, ph);
        float c = wscale;
        if (cmode == 1 && m != 0 && 2 * m != N) c *= 2.0f;
        Fn = make_float2(c * xn.x, c * xn.y);
        Fs = make_float2(c * xs.x, c * xs.y);
        if (pr.single) {
          S = Fn;
          D = S;
        } else {
          S = make_float2(Fn.x + Fs.x, Fn.y + Fs.y);
          D = make_float2(Fn.x - Fs.x, Fn.y - Fs.y);
        }
      }
      if (layout == 0) {
        float* base = X + (int64_t)m * mstride + (int64_t)k * ncol + 2 * (b0 + g);
        *reinterpret_cast<float2*>(base) = S;
        *reinterpret_cast<float2*>(base + pstride) = D;
      } else {
        float2* ftm = reinterpret_cast<float2*>(X) + (b0 + g) * nring * (int64_t)mmax;
        ftm[(int64_t)k * mmax + m] = Fn;
        if (!pr.single) ftm[(int64_t)(nring - 1 - k) * mmax + m] = Fs;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void ring_synthesis_kernel(
    const float* __restrict__ X, float* __restrict__ f, const RingPair* __restrict__ pairs,
    const float2* __restrict__ phase, const float2* __restrict__ chirp_all,
    const float2* __restrict__ bhat_all, const float2* __restrict__ tw, int log2_tab,
    int64_t npix, int64_t nbatch, int mmax, int kpad, int64_t ncol, int use_weight,
    float scale, int cmode, int layout, int nring) {
  __shared__ float2 buf[kFftLds];
  const int k = blockIdx.x;
  const RingPair pr = pairs[k];
  const int N = pr.nphi;
  const bool blue = pr.M != N;
  const float2* chirp = chirp_all + pr.chirp_off;
  const float2* bhat = bhat_all + pr.bhat_off;
  int log2G = 13 - pr.log2M;
  if (log2G > 3) log2G = 3;
  const int G = 1 << log2G;
  const int64_t b_begin = (int64_t)blockIdx.y * kFieldsPerFftBlock;
  int64_t b_end = b_begin + kFieldsPerFftBlock;
  if (b_end > nbatch) b_end = nbatch;
  const float wscale = (use_weight ? pr.weight : 1.0f) * scale;
  const int K = min(mmax, N / 2 + 1);
  const int64_t pstride = (int64_t)kpad * ncol;
  const int64_t mstride = 2 * pstride;
  const int total = G << pr.log2M;

  for (int64_t b0 = b_begin; b0 < b_end; b0 += G) {
    const int gcount = (int)min((int64_t)G, b_end - b0);
    for (int idx = threadIdx.x; idx < total; idx += kThreads) buf[idx] = make_float2(0.f, 0.f);
    __syncthreads();
    for (int e = threadIdx.x; e < (K << log2G); e += kThreads) {
      const int m = e >> log2G;
      const int g = e & (G - 1);
      if (g >= gcount) continue;
      float2 Gn, Gs;
      if (layout == 0) {
        const float* base = X + (int64_t)m * mstride + (int64_t)k * ncol + 2 * (b0 + g);
        const float2 S = *reinterpret_cast<const float2*>(base);
        const float2 D = *reinterpret_cast<const float2*>(base + pstride);
        Gn = make_float2(S.x + D.x, S.y + D.y);
        Gs = make_float2(S.x - D.x, S.y - D.y);
      } else {
        const float2* ftm = reinterpret_cast<const float2*>(X) + (b0 + g) * nring * (int64_t)mmax;
        Gn = ftm[(int64_t)k * mmax + m];
        Gs = pr.single ? Gn : ftm[(int64_t)(nring - 1 - k) * mmax + m];
      }
      const float2 ph = cconj(phase[(int64_t)k * mmax + m]);
      const bool edge = (m == 0) || (2 * m == N);
      float c = wscale;
      if (cmode == 1 && !edge) c *= 0.5f;
      float2 xn = cmul(make_float2(c * Gn.x, c * Gn.y), ph);
      float2 xs = pr.single ? make_float2(0.f, 0.f) : cmul(make_float2(c * Gs.x, c * Gs.y), ph);
      if (edge) {
        xn.y = 0.f;
        xs.y = 0.f;
      }
      // Z = Xn + i Xs at m and its Hermitian partner at N - m; the DFT below
      // is applied to conj(Z) (an inverse DFT by conjugation).
      float2 zp = make_float2(xn.x - xs.y, xn.y + xs.x);
      zp = cconj(zp);
      if (blue) zp = cmul(zp, chirp[m]);
      buf[g * pr.M + m] = zp;
      if (!edge) {
        float2 zn = make_float2(xn.x + xs.y, xs.x - xn.y);
        zn = cconj(zn);
        if (blue) zn = cmul(zn, chirp[N - m]);
        buf[g * pr.M + (N - m)] = zn;
      }
    }
    __syncthreads();
    ring_dft(buf, pr, total, bhat, tw, log2_tab);
    for (int g = 0; g < gcount; ++g) {
      float* fr = f + (b0 + g) * npix;
      for (int j = threadIdx.x; j < N; j += kThreads) {
        const float2 r = dft_value(buf, pr, g, j, chirp);
        fr[pr.off_n + j] = r.x;
        if (!pr.single) fr[pr.off_s + j] = -r.y;
      }
    }
    __syncthreads();
  }
}

}  // namespace

int launch_analysis(const ShtPlan& p, const float* f, float* X, int64_t nbatch, int use_weight,
                    float scale, int cmode, hipStream_t s, int layout) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ygroups = (nbatch + kFieldsPerFftBlock - 1) / kFieldsPerFftBlock;
  if (ygroups > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large (%lld)", (long long)nbatch);
  dim3 grid(p.npair, (unsigned)ygroups);
  ring_analysis_kernel<<<grid, kThreads, 0, s>>>(f, X, p.d_pairs, p.d_phase, p.d_chirp, p.d_bhat,
                                                 p.d_twiddle, p.log2_max_M, p.npix, nbatch, p.mmax,
                                                 p.kpad, col_pad(nbatch), use_weight, scale, cmode,
                                                 layout, p.nring);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

int launch_synthesis(const ShtPlan& p, const float* X, float* f, int64_t nbatch, int use_weight,
                     float scale, int cmode, hipStream_t s, int layout) {
  if (nbatch <= 0) return HPX_OK;
  const int64_t ygroups = (nbatch + kFieldsPerFftBlock - 1) / kFieldsPerFftBlock;
  if (ygroups > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large (%lld)", (long long)nbatch);
  dim3 grid(p.npair, (unsigned)ygroups);
  ring_synthesis_kernel<<<grid, kThreads, 0, s>>>(X, f, p.d_pairs, p.d_phase, p.d_chirp, p.d_bhat,
                                                  p.d_twiddle, p.log2_max_M, p.npix, nbatch, p.mmax,
                                                  p.kpad, col_pad(nbatch), use_weight, scale, cmode,
                                                 layout, p.nring);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

}  // namespace hpxmi
