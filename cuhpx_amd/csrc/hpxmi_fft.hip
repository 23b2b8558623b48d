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
// MI355X design.  The two real rings of a mirror pair (north t and south
// nring-1-t share length, phase and weight) are packed into one complex
// sequence z = f_t + i f_t', so one complex DFT serves both.  Each DFT is
// owned by 1, 2 or 4 wavefronts (a launch shape the autotuner picks) and runs
// in the field's slice of LDS: a radix-16/8/4/2 Stockham FFT.  With one
// wavefront per field the passes need no work-group barrier (LDS accesses of
// one wavefront complete in program order), so the waves of a work-group
// progress independently; wider fields synchronise the work-group between
// passes.  Rings whose length is not a power of two use Bluestein's chirp-z
// algorithm with a padded power-of-two length M >= 2N-1 and FP64-built
// chirp / kernel-spectrum tables.  A work-group holds one ring pair and a
// few fields; pairs are launched in classes of equal maximum FFT length so
// that the LDS allocation fits the class.  Chirp multiply, phase shift,
// quadrature weight, the split into the two rings and the equatorial folding
// S = F_t + F_t', D = F_t - F_t' (the Legendre step's operands for l+m even /
// odd) are fused into the load and store loops; the folded spectra are
// written through LDS so that consecutive lanes write consecutive fields.
// Bluestein's pointwise steps run inside the first and last Stockham passes,
// and the first pass of an analysis reads the map (the last pass of a
// synthesis writes it) directly.
// Rings that need an 8192- or 16384-point FFT (nside > 512) use a
// work-group-wide variant of the same Stockham passes (up to 128 KB of LDS).
// Even rings whose pair would need a longer FFT than LDS holds (nside 4096
// and 8192) are transformed one ring at a time as a half- or quarter-length
// complex sequence.

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

namespace hpxmi {
namespace {

// Explicit fused multiply-adds: the same rounding in every kernel shape (a
// plain a*b - c*d may be contracted either way, differently per instantiation,
// which would make results depend on the launch shape).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(s * a.x, s * a.y); }
// multiply by -i
__device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// LDS position of point x of a wave's sequence: one float2 of padding after
// every 16 points, so that the strided Stockham writes (stride R points in the
// first pass) spread over the banks.  Sequences of consecutive waves are
// wave_stride(M) float2 apart (two more), so that the same point of different
// waves falls in different banks in the cross-wave store loops.
__device__ __forceinline__ int pad(int x) { return x + (x >> 4); }
__host__ __device__ constexpr int wave_stride(int m) { return m + (m >> 4) + 2; }

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

template <>
__device__ __forceinline__ void dft<16>(float2* v) {
  float2 e[8], o[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    e[q] = v[2 * q];
    o[q] = v[2 * q + 1];
  }
  dft<8>(e);
  dft<8>(o);
  // o[k] *= exp(-2 pi i k / 16)
  const float c1 = 0.923879533f, s1 = 0.382683432f, h = 0.70710678118654752f;
  o[1] = cmul(o[1], make_float2(c1, -s1));
  o[2] = make_float2((o[2].x + o[2].y) * h, (o[2].y - o[2].x) * h);
  o[3] = cmul(o[3], make_float2(s1, -c1));
  o[4] = cmul_mi(o[4]);
  o[5] = cmul(o[5], make_float2(-s1, -c1));
  o[6] = make_float2((o[6].y - o[6].x) * h, -(o[6].x + o[6].y) * h);
  o[7] = cmul(o[7], make_float2(-c1, -s1));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    v[k] = cadd(e[k], o[k]);
    v[k + 8] = csub(e[k], o[k]);
  }
}

template <int R>
constexpr int log2_of() {
  return R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1));
}

// v[r] *= w^r for r = 1..R-1, powers of the table entry by multiplication:
// sequentially up to R = 8; for R = 16 as w^r = w^(r & (r-1)) w^(r & -r)
// from the squares w^2, w^4, w^8 (at most 6 roundings deep instead of 14).
template <int R>
__device__ __forceinline__ void twiddle(float2* v, float2 w1) {
  if constexpr (R == 16) {
    float2 pw[16];
    pw[1] = w1;
    pw[2] = cmul(w1, w1);
    pw[4] = cmul(pw[2], pw[2]);
    pw[8] = cmul(pw[4], pw[4]);
#pragma unroll
    for (int r = 3; r < 16; ++r)
      if (r & (r - 1)) pw[r] = cmul(pw[r & (r - 1)], pw[r & -r]);
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = cmul(v[r], pw[r]);
  } else {
    float2 w = w1;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      v[r] = cmul(v[r], w);
      if (r + 1 < R) w = cmul(w, w1);
    }
  }
}

// ---------------------------------------------------------------------------
// Wave-level FFT: one wavefront, one sequence of M = 2^log2M points in LDS,
// natural order in and out.
// ---------------------------------------------------------------------------

__device__ __forceinline__ void wave_sync() {
  // The LDS accesses of one wavefront are performed in program order; this
  // only keeps the compiler from moving them across a pass boundary.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Barrier between the passes of one field's DFT: the field's own wavefront
// (TPF = 64 threads per field), or the work-group when a field is spread over
// several wavefronts.
template <int TPF>
__device__ __forceinline__ void field_sync() {
  if constexpr (TPF == 64)
    wave_sync();
  else
    __syncthreads();
}

// One radix-R Stockham pass over M = 2^LOG2M points by one wavefront, after
// 2^LOG2NS points have been combined.  All strides are compile-time constants,
// so the LDS accesses use immediate offsets.
// The global-memory ends of a ring pair's DFT, used by its first and last
// passes (n: the ring length; 0 for a field beyond the batch, which then
// reads zeros and writes nothing).
struct PassIO {
  const float2* pro = nullptr;  // pre-multiply factors (chirp) or nullptr
  const float2* epi = nullptr;  // post-multiply factors
  const float* in_n = nullptr;  // ring samples, north / south ring
  const float* in_s = nullptr;
  float* out_n = nullptr;       // output rings
  float* out_s = nullptr;
  int n = 0;
  int single = 0;               // the equator: one real ring
};

// TPF threads share the sequence (lane = thread index within the field).
// Ends fused into the pass (input x = b + r kNB in the first pass, output
// x = b + r kNS in the last):
//   PRO = 1: input x *= pro[x] for x < n (Bluestein chirp pre-multiply);
//   PRO = 2: input x = (in_n[x], in_s[x]) (times pro[x] if given) for x < n,
//            0 beyond, read from global memory instead of LDS;
//   EPI = 1: output x = conj(output x * epi[x]) (Bluestein kernel spectrum);
//   EPI = 2: output x = epi[x] conj(output x) for x < n (chirp post-multiply);
//   EPI = 3: as 2, written to out_n[x] (real part), out_s[x] (minus the
//            imaginary part) instead of LDS;
//   EPI = 4: output x written to out_n / out_s likewise, for x < n.
template <int R, int LOG2M, int LOG2NS, int TPF = 64, int PRO = 0, int EPI = 0>
__device__ __forceinline__ void wave_pass(float2* buf, const float2* __restrict__ tw, int log2_tab,
                                          int lane, const PassIO& io = PassIO{}) {
  constexpr int kLogR = log2_of<R>();
  constexpr int kNB = 1 << (LOG2M - kLogR);  // butterflies; also the input stride
  constexpr int kNS = 1 << LOG2NS;
  constexpr int kB = (kNB + TPF - 1) / TPF;
  static_assert(PRO == 0 || LOG2NS == 0, "pre-multiply on the first pass only");
  static_assert(EPI == 0 || LOG2NS + kLogR == LOG2M, "post-multiply on the last pass only");
  const int tw_shift = log2_tab - LOG2NS - kLogR;
  float2 v[kB][R];
  float2 w[kB];
#pragma unroll
  for (int i = 0; i < kB; ++i) {
    const int b = lane + TPF * i;
    if (kNS > 1 && (kNB >= TPF || b < kNB)) w[i] = tw[(b & (kNS - 1)) << tw_shift];
  }
  // padded positions: for power-of-two strides s, pad(base + r s) =
  // pad(base) + r s + ((r s) >> 4) when s is a multiple of 16 (reads) or base
  // is a Stockham output base (writes), so the offsets stay immediate
#pragma unroll
  for (int i = 0; i < kB; ++i) {
    const int b = lane + TPF * i;
    if (kNB >= TPF || b < kNB) {
      if constexpr (PRO == 2) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int x = b + r * kNB;
          v[i][r] = make_float2(0.f, 0.f);
          if (x < io.n) v[i][r] = make_float2(io.in_n[x], io.single ? 0.f : io.in_s[x]);
        }
      } else {
        const int pb = pad(b);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if constexpr (kNB % 16 == 0)
            v[i][r] = buf[pb + r * kNB + ((r * kNB) >> 4)];
          else
            v[i][r] = buf[pad(b + r * kNB)];
        }
      }
    }
  }
  // the pointwise factors, requested while the reads complete
  constexpr bool kPreMul = PRO == 1 || PRO == 2;
  constexpr bool kPostMul = EPI == 1 || EPI == 2 || EPI == 3;
  float2 fpre[kPreMul ? kB : 1][kPreMul ? R : 1];
  float2 f[kPostMul ? kB : 1][kPostMul ? R : 1];
  if constexpr (kPreMul) {
    if (io.pro != nullptr) {
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int b = lane + TPF * i;
        if (kNB >= TPF || b < kNB) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (b + r * kNB < io.n) fpre[i][r] = io.pro[b + r * kNB];
        }
      }
    }
  }
  if constexpr (kPostMul) {
#pragma unroll
    for (int i = 0; i < kB; ++i) {
      const int b = lane + TPF * i;
      if (kNB >= TPF || b < kNB) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (EPI == 1 || b + r * kNS < io.n) f[i][r] = io.epi[b + r * kNS];
      }
    }
  }
  if constexpr (kPreMul) {
    if (io.pro != nullptr) {
#pragma unroll
      for (int i = 0; i < kB; ++i) {
        const int b = lane + TPF * i;
        if (kNB >= TPF || b < kNB) {
#pragma unroll
          for (int r = 0; r < R; ++r)
            if (b + r * kNB < io.n) v[i][r] = cmul(v[i][r], fpre[i][r]);
        }
      }
    }
  }
  field_sync<TPF>();
#pragma unroll
  for (int i = 0; i < kB; ++i) {
    const int b = lane + TPF * i;
    if (kNB >= TPF || b < kNB) {
      const int jn = b & (kNS - 1);
      if (kNS > 1) twiddle<R>(v[i], w[i]);
      dft<R>(v[i]);
      if constexpr (EPI == 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[i][r] = cconj(cmul(v[i][r], f[i][r]));
      } else if constexpr (EPI == 2 || EPI == 3) {
#pragma unroll
        for (int r = 0; r < R; ++r)
          if (b + r * kNS < io.n) v[i][r] = cmul(f[i][r], cconj(v[i][r]));
      }
      if constexpr (EPI == 3 || EPI == 4) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int x = b + r * kNS;
          if (x < io.n) {
            io.out_n[x] = v[i][r].x;
            if (!io.single) io.out_s[x] = -v[i][r].y;
          }
        }
      } else {
        const int pbase = pad((b - jn) * R + jn);
#pragma unroll
        for (int r = 0; r < R; ++r) buf[pbase + r * kNS + ((r * kNS) >> 4)] = v[i][r];
      }
    }
  }
  field_sync<TPF>();
}

// Radix schedule: radix 16 while at least 6 bits remain or exactly 4, then
// 8 for 3 or 5 remaining bits, 4 for 2, 2 for 1 (1024 and 2048 points in 3
// passes, 4096 in 3).
template <int LOG2M, int LOG2NS = 0, int TPF = 64, int PRO = 0, int EPI = 0>
__device__ __forceinline__ void wave_fft_fixed(float2* buf, const float2* __restrict__ tw,
                                               int log2_tab, int lane, const PassIO& io = PassIO{}) {
  if constexpr (LOG2NS < LOG2M) {
    constexpr int rem = LOG2M - LOG2NS;
    constexpr int R = (rem >= 6 || rem == 4) ? 16 : ((rem == 3 || rem == 5) ? 8 : (rem == 2 ? 4 : 2));
    constexpr bool first = LOG2NS == 0, last = LOG2NS + log2_of<R>() == LOG2M;
    wave_pass<R, LOG2M, LOG2NS, TPF, first ? PRO : 0, last ? EPI : 0>(buf, tw, log2_tab, lane, io);
    wave_fft_fixed<LOG2M, LOG2NS + log2_of<R>(), TPF, PRO, EPI>(buf, tw, log2_tab, lane, io);
  }
}

// Forward DFT of 2^log2M points (natural order in and out) by one wavefront,
// for the sizes of one size class (log2M in [LO, HI]).
template <int LO, int HI, int TPF = 64, int PRO = 0, int EPI = 0>
__device__ __forceinline__ void wave_fft(float2* buf, int log2M, const float2* __restrict__ tw,
                                         int log2_tab, int lane, const PassIO& io = PassIO{}) {
  if constexpr (LO <= HI) {
    if (log2M == LO) {
      wave_fft_fixed<LO, 0, TPF, PRO, EPI>(buf, tw, log2_tab, lane, io);
      return;
    }
    wave_fft<LO + 1, HI, TPF, PRO, EPI>(buf, log2M, tw, log2_tab, lane, io);
  }
}

// DFT of length N of a ring pair's packed sequence z_j = f_n[j] + i f_s[j]
// (zero-padded to M for Bluestein, with X[x] = chirp_x conj(c_x),
// c = FFT(conj(FFT(z chirp) bhat)), bhat = FFT(conj chirp) / M; the pointwise
// steps run inside the first and last passes of the two FFTs).
//   ANALYSIS: z is read from io.in_n / io.in_s; afterwards point x < N of the
//     wave's LDS sequence holds the DFT value at x.
//   synthesis: z is the LDS sequence; the real and minus the imaginary parts
//     of the DFT values x < N go to io.out_n / io.out_s.
template <int LO, int HI, int TPF, bool ANALYSIS>
__device__ __forceinline__ void wave_ring_dft(float2* buf, const RingPair& pr,
                                              const float2* __restrict__ bhat,
                                              const float2* __restrict__ chirp,
                                              const float2* __restrict__ tw, int log2_tab, int lane, PassIO io) {
  if (pr.M == pr.nphi) {
    io.pro = nullptr;
    wave_fft<LO, HI, TPF, ANALYSIS ? 2 : 0, ANALYSIS ? 0 : 4>(buf, pr.log2M, tw, log2_tab, lane, io);
    return;
  }
  io.pro = chirp;
  io.epi = bhat;
  wave_fft<LO, HI, TPF, ANALYSIS ? 2 : 1, 1>(buf, pr.log2M, tw, log2_tab, lane, io);
  io.epi = chirp;
  wave_fft<LO, HI, TPF, 0, ANALYSIS ? 2 : 3>(buf, pr.log2M, tw, log2_tab, lane, io);
}

template <bool PADDED>
__device__ __forceinline__ float2 dft_value(const float2* seq, const RingPair& pr, int x,
                                            const float2* __restrict__ chirp) {
  const float2 v = seq[PADDED ? pad(x) : x];
  if (pr.M == pr.nphi) return v;
  return cmul(chirp[x], cconj(v));
}

// The two rings' spectra at m from the packed DFT Z (zm = Z[m], zc = Z[N-m]):
// Fn = (Z[m] + conj Z[N-m]) / 2, Fs = (Z[m] - conj Z[N-m]) / (2i), each times
// the phase factor and the scale c.
__device__ __forceinline__ void unpack_pair(float2 zm, float2 zc, float2 ph, float c, float2& Fn,
                                            float2& Fs) {
  const float2 xn = make_float2(0.5f * (zm.x + zc.x), 0.5f * (zm.y - zc.y));
  const float2 xs = make_float2(0.5f * (zm.y + zc.y), 0.5f * (zc.x - zm.x));
  Fn = cscale(cmul(xn, ph), c);
  Fs = cscale(cmul(xs, ph), c);
}

// Synthesis input at m for the two rings' spectra Gn, Gs: the DFT below is
// applied to conj(Z) with Z = Xn + i Xs Hermitian-extended; writes positions
// m and (unless m is DC or Nyquist) N - m of the sequence, chirp-multiplied
// for Bluestein when chirp is given (chirp == nullptr: the first Stockham
// pass applies it instead).
template <bool PADDED>
__device__ __forceinline__ void pack_pair(float2* seq, const RingPair& pr, int m, float2 Gn,
                                          float2 Gs, float2 ph_conj, float c, int cmode,
                                          const float2* __restrict__ chirp) {
  const int N = pr.nphi;
  const bool edge = (m == 0) || (2 * m == N);
  if (cmode == 1 && !edge) c *= 0.5f;
  float2 xn = cmul(cscale(Gn, c), ph_conj);
  float2 xs = pr.single ? make_float2(0.f, 0.f) : cmul(cscale(Gs, c), ph_conj);
  if (edge) {
    xn.y = 0.f;
    xs.y = 0.f;
  }
  const bool blue = pr.M != N && chirp != nullptr;
  float2 zp = cconj(make_float2(xn.x - xs.y, xn.y + xs.x));
  if (blue) zp = cmul(zp, chirp[m]);
  seq[PADDED ? pad(m) : m] = zp;
  if (!edge) {
    float2 zn = cconj(make_float2(xn.x + xs.y, xs.x - xn.y));
    if (blue) zn = cmul(zn, chirp[N - m]);
    seq[PADDED ? pad(N - m) : N - m] = zn;
  }
}

struct StageArgs {
  const RingPair* pairs;
  const int* plist;       // pair indices of this class
  const float2* phase;    // [npair][mmax]
  const float2* chirp;
  const float2* bhat;
  const float2* tw;
  int log2_tab;
  int64_t npix, nbatch;
  int mmax, kpad, nring;
  int64_t ncol;
  int use_weight, cmode, layout;
  float scale;
};

// ---------------------------------------------------------------------------
// Wave-level kernels: work-group = one ring pair x WAVES fields
// ---------------------------------------------------------------------------

// Spectra of the WAVES fields of a work-group, in the LDS sequences
// (lds + g * stride, point x at pad(x); the ring pair's DFT values), to the
// folded spectra X[m][parity][k][2 b + re/im] (layout 0, all threads, fields
// fastest so that consecutive lanes write consecutive columns) or to the
// ring-major ftm (layout 1, one wave per field).
template <int WAVES, int TPF = 64>
__device__ __forceinline__ void analysis_store(const StageArgs& a, const RingPair& pr, int k, int64_t b0,
                                               int64_t b, const float2* lds, int stride, const float2* phs,
                                               float* __restrict__ X, int lane) {
  // phs[m] = exp(-i m phi0) c_m / 2, zero from K on.  With u = Z[m] and
  // v = conj Z[N-m], the rings' spectra are Fn = phs (u + v) and
  // Fs = phs (-i) (u - v), so S = Fn + Fs = phs (p - i q) and
  // D = Fn - Fs = phs (p + i q) with p = u + v, q = u - v.
  constexpr int kThreadsW = WAVES * TPF;
  const int N = pr.nphi;
  const float2* buf = lds + (threadIdx.x / TPF) * stride;
  const int K = min(a.mmax, N / 2 + 1);
  if (a.layout == 0) {
    // folded spectra X[m][parity][k][2b + re/im]: fields fastest so that
    // consecutive lanes write consecutive columns
    const int64_t pstride = (int64_t)a.kpad * a.ncol;
    const int total = pr.mwrite * WAVES;  // no further order of this pair is read
    for (int e = threadIdx.x; e < total; e += kThreadsW) {
      const int m = e / WAVES;
      const int g = e - m * WAVES;
      if (b0 + g >= a.nbatch) continue;
      const float2* seq = lds + g * stride;
      const int mi = min(m, K - 1);  // past K the phase factor is zero
      const float2 u = seq[pad(mi)];
      const float2 zc = seq[pad(mi == 0 ? 0 : N - mi)];
      const float2 ph = phs[m];
      const float2 p = make_float2(u.x + zc.x, u.y - zc.y);
      float2 S, D;
      if (pr.single) {
        S = cmul(p, ph);
        D = S;
      } else {
        const float2 q = make_float2(u.x - zc.x, u.y + zc.y);
        S = cmul(make_float2(p.x + q.y, p.y - q.x), ph);
        D = cmul(make_float2(p.x - q.y, p.y + q.x), ph);
      }
      float* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * (b0 + g);
      *reinterpret_cast<float2*>(base) = S;
      *reinterpret_cast<float2*>(base + pstride) = D;
    }
  } else if (b < a.nbatch) {
    // ring-major ftm[b][t][m]: the field's threads along m
    float2* out = reinterpret_cast<float2*>(X) + b * a.nring * (int64_t)a.mmax;
    for (int m = lane; m < a.mmax; m += TPF) {
      const int mi = min(m, K - 1);
      const float2 u = buf[pad(mi)];
      const float2 zc = buf[pad(mi == 0 ? 0 : N - mi)];
      const float2 ph = phs[m];
      const float2 Fn = cmul(make_float2(u.x + zc.x, u.y - zc.y), ph);
      const float2 Fs = cmul(make_float2(u.y + zc.y, zc.x - u.x), ph);
      out[(int64_t)k * a.mmax + m] = Fn;
      if (!pr.single) out[(int64_t)(a.nring - 1 - k) * a.mmax + m] = Fs;
    }
  }
}

// WAVES fields per work-group, TPF threads (TPF / 64 wavefronts) per field.
template <int WAVES, int LO, int HI, int TPF = 64>
__global__ __launch_bounds__(WAVES * TPF) void ring_analysis_wave_kernel(const float* __restrict__ f,
                                                                         float* __restrict__ X,
                                                                         StageArgs a) {
  extern __shared__ float2 lds[];
  constexpr int kThreadsW = WAVES * TPF;
  const int64_t nfg = (a.nbatch + WAVES - 1) / WAVES;
  const int64_t item = xcd_local_index(blockIdx.x, gridDim.x);
  const int k = a.plist[item / nfg];
  const int64_t b0 = (item % nfg) * WAVES;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const float2* chirp = a.chirp + pr.chirp_off;
  const int wave = threadIdx.x / TPF, lane = threadIdx.x % TPF;
  const int64_t b = b0 + wave;
  const int stride = wave_stride(M);
  float2* buf = lds + wave * stride;
  // the pair's phase factors exp(-i m phi0) times c_m / 2 (c_m: the
  // quadrature weight and scale, doubled off DC and Nyquist in cmode 1;
  // zero from K on), staged once for the store loop
  float2* phs = lds + WAVES * wave_stride(1 << HI);
  {
    const float2* ph = a.phase + (int64_t)k * a.mmax;
    const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
    const int K = min(a.mmax, N / 2 + 1);
    for (int m = threadIdx.x; m < a.mmax; m += kThreadsW) {
      const float c = (a.cmode == 1 && m != 0 && 2 * m != N) ? 2.0f * wscale : wscale;
      phs[m] = m < K ? cscale(ph[m], 0.5f * c) : make_float2(0.f, 0.f);
    }
  }
  // with several wavefronts per field the DFT has work-group barriers, so
  // every field runs it (fields beyond the batch on zeros); the first pass
  // reads the ring samples straight from the map
  const bool active = b < a.nbatch;
  if (active || TPF > 64) {
    PassIO io;
    io.in_n = f + (active ? b : 0) * a.npix + pr.off_n;
    io.in_s = f + (active ? b : 0) * a.npix + (pr.single ? pr.off_n : pr.off_s);
    io.n = active ? N : 0;
    io.single = pr.single;
    wave_ring_dft<LO, HI, TPF, true>(buf, pr, a.bhat + pr.bhat_off, chirp, a.tw, a.log2_tab, lane, io);
  }
  __syncthreads();
  analysis_store<WAVES, TPF>(a, pr, k, b0, b, lds, stride, phs, X, lane);
}

template <int WAVES, int LO, int HI, int TPF = 64>
__global__ __launch_bounds__(WAVES * TPF) void ring_synthesis_wave_kernel(const float* __restrict__ X,
                                                                          float* __restrict__ f,
                                                                          StageArgs a) {
  extern __shared__ float2 lds[];
  constexpr int kThreadsW = WAVES * TPF;
  const int64_t nfg = (a.nbatch + WAVES - 1) / WAVES;
  const int64_t item = xcd_local_index(blockIdx.x, gridDim.x);
  const int k = a.plist[item / nfg];
  const int64_t b0 = (item % nfg) * WAVES;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const float2* chirp = a.chirp + pr.chirp_off;
  const int wave = threadIdx.x / TPF, lane = threadIdx.x % TPF;
  const int64_t b = b0 + wave;
  const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  // folded spectra (S, D) of element e = (m = e / WAVES, field g = e % WAVES),
  // kBatch elements per thread per batch; the first batch is requested
  // before the LDS set-up and each next batch while the current one is
  // packed.  One element per batch measured fastest (ring synthesis at nside
  // 256, 64 fields: 190 us with 4, 185 with 2, 177 with 1; 219 with 8).
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  const int total = K * WAVES;
  constexpr int kBatch = 1;
  float2 S[kBatch], D[kBatch];
  auto load_batch = [&](int e0) {
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int e = e0 + u * kThreadsW;
      const int m = e / WAVES;
      const int g = e - m * WAVES;
      if (e < total && b0 + g < a.nbatch) {
        const float* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * (b0 + g);
        S[u] = *reinterpret_cast<const float2*>(base);
        D[u] = *reinterpret_cast<const float2*>(base + pstride);
      }
    }
  };
  // the pair's phase factors exp(-i m phi0), staged once for the packing.
  // They are requested before the spectra: storing them to LDS waits for
  // their loads, which would otherwise also wait for every spectra load
  // issued before them (ring synthesis at nside 256, 64 fields: 184 -> 179 us)
  float2* phs = lds + WAVES * wave_stride(1 << HI);
  {
    const float2* ph = a.phase + (int64_t)k * a.mmax;
    for (int m = threadIdx.x; m < a.mmax; m += kThreadsW) phs[m] = ph[m];
  }
  if (a.layout == 0) load_batch(threadIdx.x);

  const int stride = wave_stride(M);
  // the packing writes points m and N - m for m < K; when M == N and
  // K = N/2 + 1 that is every point, and the sequences need no zero fill
  if (M != N || 2 * (K - 1) < N)
    for (int idx = threadIdx.x; idx < WAVES * stride; idx += kThreadsW) lds[idx] = make_float2(0.f, 0.f);
  __syncthreads();
  // the packing leaves the chirp factors to a per-wave pass below
  if (a.layout == 0) {
    for (int e0 = threadIdx.x; e0 < total; e0 += kBatch * kThreadsW) {
      float2 Sc[kBatch], Dc[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        Sc[u] = S[u];
        Dc[u] = D[u];
      }
      if (e0 + kBatch * kThreadsW < total) load_batch(e0 + kBatch * kThreadsW);
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int e = e0 + u * kThreadsW;
        const int m = e / WAVES;
        const int g = e - m * WAVES;
        if (e < total && b0 + g < a.nbatch)
          pack_pair<true>(lds + g * stride, pr, m, cadd(Sc[u], Dc[u]), csub(Sc[u], Dc[u]), cconj(phs[m]),
                          wscale, a.cmode, nullptr);
      }
    }
  } else {
    for (int e = threadIdx.x; e < K * WAVES; e += kThreadsW) {
      const int g = e / K;
      const int m = e - g * K;
      if (b0 + g >= a.nbatch) continue;
      const float2* in = reinterpret_cast<const float2*>(X) + (b0 + g) * a.nring * (int64_t)a.mmax;
      const float2 Gn = in[(int64_t)k * a.mmax + m];
      const float2 Gs = pr.single ? Gn : in[(int64_t)(a.nring - 1 - k) * a.mmax + m];
      pack_pair<true>(lds + g * stride, pr, m, Gn, Gs, cconj(phs[m]), wscale, a.cmode, nullptr);
    }
  }
  __syncthreads();
  // with several wavefronts per field the DFT has work-group barriers, so
  // every field runs it (fields beyond the batch on zeros)
  if (TPF == 64 && b >= a.nbatch) return;
  float2* buf = lds + wave * stride;
  // the last pass writes the two rings of the map
  PassIO io;
  io.out_n = f + (b < a.nbatch ? b : 0) * a.npix + pr.off_n;
  io.out_s = f + (b < a.nbatch ? b : 0) * a.npix + pr.off_s;
  io.n = b < a.nbatch ? N : 0;
  io.single = pr.single;
  wave_ring_dft<LO, HI, TPF, false>(buf, pr, a.bhat + pr.bhat_off, chirp, a.tw, a.log2_tab, lane, io);
}

// ---------------------------------------------------------------------------
// Work-group-wide FFT for the 8192- and 16384-point class: one ring pair x
// one field per work-group, Stockham passes across 512 threads with barriers.
// ---------------------------------------------------------------------------

constexpr int kBlockThreads = 512;

template <int R>
__device__ __forceinline__ void block_pass(float2* buf, int log2M, int Ns, int log2Ns,
                                           const float2* __restrict__ tw, int log2_tab) {
  constexpr int kLogR = log2_of<R>();
  constexpr int kMaxB = (kFftLds / R + kBlockThreads - 1) / kBlockThreads;
  const int log2mr = log2M - kLogR;
  const int nb = 1 << log2mr;
  const int tw_shift = log2_tab - log2Ns - kLogR;
  float2 v[kMaxB][R];
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kBlockThreads;
    if (b < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[b + (r << log2mr)];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < kMaxB; ++i) {
    const int b = threadIdx.x + i * kBlockThreads;
    if (b < nb) {
      const int jn = b & (Ns - 1);
      if (Ns > 1) twiddle<R>(v[i], tw[jn << tw_shift]);
      dft<R>(v[i]);
      float2* dst = buf + (b - jn) * R + jn;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * Ns] = v[i][r];
    }
  }
  __syncthreads();
}

__device__ void block_fft(float2* buf, int log2M, const float2* __restrict__ tw, int log2_tab) {
  int Ns = 1, log2Ns = 0, rem = log2M;
  while (rem > 0) {
    if (rem == 1) {
      block_pass<2>(buf, log2M, Ns, log2Ns, tw, log2_tab);
      rem -= 1;
      Ns <<= 1;
      log2Ns += 1;
    } else if (rem == 2 || rem == 4) {
      block_pass<4>(buf, log2M, Ns, log2Ns, tw, log2_tab);
      rem -= 2;
      Ns <<= 2;
      log2Ns += 2;
    } else {
      block_pass<8>(buf, log2M, Ns, log2Ns, tw, log2_tab);
      rem -= 3;
      Ns <<= 3;
      log2Ns += 3;
    }
  }
}

__device__ void block_ring_dft(float2* buf, const RingPair& pr, const float2* __restrict__ bhat,
                               const float2* __restrict__ tw, int log2_tab) {
  block_fft(buf, pr.log2M, tw, log2_tab);
  if (pr.M != pr.nseq) {
    for (int x = threadIdx.x; x < pr.M; x += kBlockThreads) buf[x] = cconj(cmul(buf[x], bhat[x]));
    __syncthreads();
    block_fft(buf, pr.log2M, tw, log2_tab);
  }
}

// Rings transformed one at a time (RingPair::half: the pair's packed
// sequence would need a longer FFT than LDS holds).  The ring's N samples
// become H = N/2 complex points z_j = f_2j + i f_2j+1.  With Z = DFT_H(z), the
// real ring's DFT is
//   F[m] = E[m] + e^{-2 pi i m / N} O[m],  E = (Z[m] + conj Z[H-m]) / 2,
//   O = (Z[m] - conj Z[H-m]) / (2i)   (indices mod H; m <= H),
// and its inverse packs Y[m] = E'[m] + i O'[m] with E' = F[m] + F[m+H],
// O' = (F[m] - F[m+H]) e^{2 pi i m / N}, F[m+H] = conj F[H-m], so that
// f_2n + i f_2n+1 = IDFT_H(Y)[n] = conj(DFT_H(conj Y))[n].
// half = 1: DFT_H in one transform (Bluestein or a power of two).
// half = 2: DFT_H by one radix-2 step over two Bluestein DFTs of Q = H/2
// points, A of the even and B of the odd points:
//   DFT_H(u)[x] = A[x mod Q] + e^{-2 pi i x / H} B[x mod Q];
// the first is kept in registers (Q <= kFftLds / 2 points over the
// work-group) while the second runs in LDS.
constexpr int kHalfRegs = kFftLds / 2 / kBlockThreads;

// DFT value x of the sequence the last block_ring_dft transformed.
__device__ __forceinline__ float2 seq_value(const float2* buf, const RingPair& pr, int x,
                                            const float2* __restrict__ chirp) {
  const float2 v = buf[x];
  return pr.M == pr.nseq ? v : cmul(chirp[x], cconj(v));
}

// In LDS: sequence j < n of a ring transform, from load(j) (n <= M; zero
// padded to M), times the chirp for Bluestein; then the transform.
template <class Load>
__device__ __forceinline__ void ring_seq_dft(float2* buf, const StageArgs& a, const RingPair& pr,
                                             const float2* __restrict__ chirp, Load load) {
  const bool blue = pr.M != pr.nseq;
  for (int j = threadIdx.x; j < pr.M; j += kBlockThreads) {
    float2 z = make_float2(0.f, 0.f);
    if (j < pr.nseq) {
      z = load(j);
      if (blue) z = cmul(z, chirp[j]);
    }
    buf[j] = z;
  }
  __syncthreads();
  block_ring_dft(buf, pr, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
}

__device__ void analysis_half(float2* buf, const StageArgs& a, const RingPair& pr, int k, int64_t b,
                              const float* __restrict__ fr, float* __restrict__ X) {
  const int N = pr.nphi, H = N >> 1, Q = pr.nseq;
  const float2* chirp = a.chirp + pr.chirp_off;
  const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const float2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int side = 0; side < (pr.single ? 1 : 2); ++side) {
    const float* src = fr + (side ? pr.off_s : pr.off_n);
    if (pr.half == 1) {
      ring_seq_dft(buf, a, pr, chirp, [&](int j) { return make_float2(src[2 * j], src[2 * j + 1]); });
    } else {
      // A = DFT_Q(z_0, z_2, ...), kept in registers; B = DFT_Q(z_1, z_3, ...)
      // in LDS at x < Q, and A moved to Q + x
      float2 Areg[kHalfRegs];
      ring_seq_dft(buf, a, pr, chirp, [&](int j) { return make_float2(src[4 * j], src[4 * j + 1]); });
#pragma unroll
      for (int i = 0; i < kHalfRegs; ++i) {
        const int x = threadIdx.x + i * kBlockThreads;
        if (x < Q) Areg[i] = seq_value(buf, pr, x, chirp);
      }
      __syncthreads();
      ring_seq_dft(buf, a, pr, chirp, [&](int j) { return make_float2(src[4 * j + 2], src[4 * j + 3]); });
#pragma unroll
      for (int i = 0; i < kHalfRegs; ++i) {
        const int x = threadIdx.x + i * kBlockThreads;
        if (x < Q) {
          buf[x] = seq_value(buf, pr, x, chirp);
          buf[Q + x] = Areg[i];
        }
      }
      __syncthreads();
    }
    // Z[x] for 0 <= x < H
    auto zval = [&](int x) -> float2 {
      if (pr.half == 1) return seq_value(buf, pr, x, chirp);
      const int xq = x < Q ? x : x - Q;
      float sn, cs;
      sincospif(-2.0f * (float)x / (float)H, &sn, &cs);
      return cadd(buf[Q + xq], cmul(make_float2(cs, sn), buf[xq]));
    };
    for (int m = threadIdx.x; m < a.mmax; m += kBlockThreads) {
      float2 F = make_float2(0.f, 0.f);
      if (m < K) {
        const float2 g1 = zval(m == H ? 0 : m);
        const float2 g2 = zval(m == 0 ? 0 : H - m);
        const float2 E = make_float2(0.5f * (g1.x + g2.x), 0.5f * (g1.y - g2.y));
        const float2 O = make_float2(0.5f * (g1.y + g2.y), -0.5f * (g1.x - g2.x));
        float sn, cs;
        sincospif(-2.0f * (float)m / (float)N, &sn, &cs);
        F = cadd(E, cmul(make_float2(cs, sn), O));
        const float c = (a.cmode == 1 && m != 0 && 2 * m != N) ? 2.0f * wscale : wscale;
        F = cscale(cmul(F, ph[m]), c);
      }
      if (a.layout == 0) {
        // the north ring's spectrum is parked in S and D, then folded with
        // the south ring's by the same thread
        float2* S = reinterpret_cast<float2*>(X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b);
        float2* D = reinterpret_cast<float2*>(reinterpret_cast<float*>(S) + pstride);
        if (side == 0) {
          *S = F;
          *D = F;
        } else {
          const float2 Fn = *S;
          *S = cadd(Fn, F);
          *D = csub(Fn, F);
        }
      } else {
        float2* out = reinterpret_cast<float2*>(X) + b * a.nring * (int64_t)a.mmax;
        out[(int64_t)(side ? a.nring - 1 - k : k) * a.mmax + m] = F;
      }
    }
    __syncthreads();
  }
}

__device__ void synthesis_half(float2* buf, const StageArgs& a, const RingPair& pr, int k, int64_t b,
                               const float* __restrict__ X, float* __restrict__ fr) {
  const int N = pr.nphi, H = N >> 1, Q = pr.nseq;
  const float2* chirp = a.chirp + pr.chirp_off;
  const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const float2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int side = 0; side < (pr.single ? 1 : 2); ++side) {
    // F[m] of this ring for 0 <= m <= H (zero from K on), as the pair
    // packing forms it: weight and phase applied, halved off DC and Nyquist
    // in cmode 1, DC and Nyquist real
    auto spec = [&](int m) -> float2 {
      if (m >= K) return make_float2(0.f, 0.f);
      float2 G;
      if (a.layout == 0) {
        const float* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
        const float2 S = *reinterpret_cast<const float2*>(base);
        const float2 D = *reinterpret_cast<const float2*>(base + pstride);
        G = (side == 0 || pr.single) ? cadd(S, D) : csub(S, D);
      } else {
        const float2* in = reinterpret_cast<const float2*>(X) + b * a.nring * (int64_t)a.mmax;
        G = in[(int64_t)(side ? a.nring - 1 - k : k) * a.mmax + m];
      }
      const bool edge = (m == 0) || (2 * m == N);
      float c = wscale;
      if (a.cmode == 1 && !edge) c *= 0.5f;
      float2 x = cmul(cscale(G, c), cconj(ph[m]));
      if (edge) x.y = 0.f;
      return x;
    };
    // u[j] = conj Y[j], j < H
    auto uval = [&](int j) -> float2 {
      const float2 f1 = spec(j);
      const float2 f2 = cconj(spec(H - j));  // F[j + H]
      const float2 Ep = cadd(f1, f2);
      float sn, cs;
      sincospif(2.0f * (float)j / (float)N, &sn, &cs);
      const float2 Op = cmul(csub(f1, f2), make_float2(cs, sn));
      return cconj(make_float2(Ep.x - Op.y, Ep.y + Op.x));
    };
    float* dst = fr + (side ? pr.off_s : pr.off_n);
    if (pr.half == 1) {
      ring_seq_dft(buf, a, pr, chirp, uval);
      for (int j = threadIdx.x; j < H; j += kBlockThreads) {
        const float2 r = seq_value(buf, pr, j, chirp);
        dst[2 * j] = r.x;
        dst[2 * j + 1] = -r.y;
      }
    } else {
      // DFT_H(u)[x] = A[x] + w^x B[x] and [x + Q] = A[x] - w^x B[x] (x < Q,
      // w = e^{-2 pi i / H}), A, B the DFTs of u's even and odd points
      float2 Areg[kHalfRegs];
      ring_seq_dft(buf, a, pr, chirp, [&](int j) { return uval(2 * j); });
#pragma unroll
      for (int i = 0; i < kHalfRegs; ++i) {
        const int x = threadIdx.x + i * kBlockThreads;
        if (x < Q) Areg[i] = seq_value(buf, pr, x, chirp);
      }
      __syncthreads();
      ring_seq_dft(buf, a, pr, chirp, [&](int j) { return uval(2 * j + 1); });
#pragma unroll
      for (int i = 0; i < kHalfRegs; ++i) {
        const int x = threadIdx.x + i * kBlockThreads;
        if (x < Q) {
          float sn, cs;
          sincospif(-2.0f * (float)x / (float)H, &sn, &cs);
          const float2 wb = cmul(make_float2(cs, sn), seq_value(buf, pr, x, chirp));
          const float2 r0 = cadd(Areg[i], wb), r1 = csub(Areg[i], wb);
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

__global__ __launch_bounds__(kBlockThreads) void ring_analysis_block_kernel(
    const float* __restrict__ f, float* __restrict__ X, StageArgs a) {
  extern __shared__ float2 buf[];  // the class's longest FFT
  const int64_t item = blockIdx.x;
  const int k = a.plist[item / a.nbatch];
  const int64_t b = item % a.nbatch;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const bool blue = M != N;
  const float2* chirp = a.chirp + pr.chirp_off;
  const float* fr = f + b * a.npix;
  if (pr.half) {
    analysis_half(buf, a, pr, k, b, fr, X);
    return;
  }
  for (int j = threadIdx.x; j < M; j += kBlockThreads) {
    float2 z = make_float2(0.f, 0.f);
    if (j < N) {
      z.x = fr[pr.off_n + j];
      z.y = pr.single ? 0.f : fr[pr.off_s + j];
      if (blue) z = cmul(z, chirp[j]);
    }
    buf[j] = z;
  }
  __syncthreads();
  block_ring_dft(buf, pr, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
  const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const float2* ph = a.phase + (int64_t)k * a.mmax;
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int m = threadIdx.x; m < a.mmax; m += kBlockThreads) {
    float2 Fn = make_float2(0.f, 0.f), Fs = Fn;
    if (m < K) {
      const float c = (a.cmode == 1 && m != 0 && 2 * m != N) ? 2.0f * wscale : wscale;
      unpack_pair(dft_value<false>(buf, pr, m, chirp), dft_value<false>(buf, pr, m == 0 ? 0 : N - m, chirp),
                  ph[m], c, Fn, Fs);
    }
    if (a.layout == 0) {
      const float2 S = pr.single ? Fn : cadd(Fn, Fs);
      const float2 D = pr.single ? Fn : csub(Fn, Fs);
      float* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
      *reinterpret_cast<float2*>(base) = S;
      *reinterpret_cast<float2*>(base + pstride) = D;
    } else {
      float2* out = reinterpret_cast<float2*>(X) + b * a.nring * (int64_t)a.mmax;
      out[(int64_t)k * a.mmax + m] = Fn;
      if (!pr.single) out[(int64_t)(a.nring - 1 - k) * a.mmax + m] = Fs;
    }
  }
}

__global__ __launch_bounds__(kBlockThreads) void ring_synthesis_block_kernel(
    const float* __restrict__ X, float* __restrict__ f, StageArgs a) {
  extern __shared__ float2 buf[];  // the class's longest FFT
  const int64_t item = blockIdx.x;
  const int k = a.plist[item / a.nbatch];
  const int64_t b = item % a.nbatch;
  const RingPair pr = a.pairs[k];
  const int N = pr.nphi, M = pr.M;
  const float2* chirp = a.chirp + pr.chirp_off;
  const float wscale = (a.use_weight ? pr.weight : 1.0f) * a.scale;
  const int K = min(a.mmax, N / 2 + 1);
  const float2* ph = a.phase + (int64_t)k * a.mmax;
  if (pr.half) {
    synthesis_half(buf, a, pr, k, b, X, f + b * a.npix);
    return;
  }
  for (int idx = threadIdx.x; idx < M; idx += kBlockThreads) buf[idx] = make_float2(0.f, 0.f);
  __syncthreads();
  const int64_t pstride = (int64_t)a.kpad * a.ncol;
  for (int m = threadIdx.x; m < K; m += kBlockThreads) {
    float2 Gn, Gs;
    if (a.layout == 0) {
      const float* base = X + (int64_t)m * 2 * pstride + (int64_t)k * a.ncol + 2 * b;
      const float2 S = *reinterpret_cast<const float2*>(base);
      const float2 D = *reinterpret_cast<const float2*>(base + pstride);
      Gn = cadd(S, D);
      Gs = csub(S, D);
    } else {
      const float2* in = reinterpret_cast<const float2*>(X) + b * a.nring * (int64_t)a.mmax;
      Gn = in[(int64_t)k * a.mmax + m];
      Gs = pr.single ? Gn : in[(int64_t)(a.nring - 1 - k) * a.mmax + m];
    }
    pack_pair<false>(buf, pr, m, Gn, Gs, cconj(ph[m]), wscale, a.cmode, chirp);
  }
  __syncthreads();
  block_ring_dft(buf, pr, a.bhat + pr.bhat_off, a.tw, a.log2_tab);
  float* fr = f + b * a.npix;
  for (int j = threadIdx.x; j < N; j += kBlockThreads) {
    const float2 r = dft_value<false>(buf, pr, j, chirp);
    fr[pr.off_n + j] = r.x;
    if (!pr.single) fr[pr.off_s + j] = -r.y;
  }
}

// Rows k in [npair, kpad) of the folded spectra (padding of the Legendre
// GEMM's reduction dimension) are zero, so that the GEMM reads them unguarded.
__global__ __launch_bounds__(256) void zero_pad_rows_kernel(float* __restrict__ X, int kpad, int npair,
                                                            int64_t ncol) {
  float* base = X + ((int64_t)blockIdx.x * kpad + npair) * ncol;
  const int64_t n = (int64_t)(kpad - npair) * ncol;
  for (int64_t i = threadIdx.x; i < n; i += 256) base[i] = 0.f;
}

StageArgs make_args(const ShtPlan& p, int64_t nbatch, int use_weight, float scale, int cmode,
                    int layout) {
  StageArgs a;
  a.pairs = p.d_pairs;
  a.plist = p.d_plist;
  a.phase = p.d_phase;
  a.chirp = p.d_chirp;
  a.bhat = p.d_bhat;
  a.tw = p.d_twiddle;
  a.log2_tab = p.log2_max_M;
  a.npix = p.npix;
  a.nbatch = nbatch;
  a.mmax = p.mmax;
  a.kpad = p.kpad;
  a.nring = p.nring;
  a.ncol = col_pad(nbatch);
  a.use_weight = use_weight;
  a.cmode = cmode;
  a.layout = layout;
  a.scale = scale;
  return a;
}

template <int WAVES, int C, int TPF = 64>
void launch_wave(bool analysis, unsigned grid, size_t lds_bytes, const float* in, float* out,
                 const StageArgs& a, hipStream_t s) {
  constexpr int kLog2 = 10 + C;
  constexpr int kLo = C == 0 ? 1 : kLog2;
  auto ka = ring_analysis_wave_kernel<WAVES, kLo, kLog2, TPF>;
  auto ks = ring_synthesis_wave_kernel<WAVES, kLo, kLog2, TPF>;
  // the padded sequences take a little more than 64 KB per work-group
  // (gfx950 allows 160 KB)
  static const bool attr_set = [&] {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ka), hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ks), hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    return true;
  }();
  (void)attr_set;
  if (analysis)
    ka<<<grid, WAVES * TPF, lds_bytes, s>>>(in, out, a);
  else
    ks<<<grid, WAVES * TPF, lds_bytes, s>>>(in, out, a);
}

}  // namespace

size_t fft_lds_bytes(const ShtPlan& p, int c, int waves) {
  return ((size_t)waves * wave_stride(1 << (10 + c)) + p.mmax) * sizeof(float2);
}

int launch_fft_class(const ShtPlan& p, bool analysis, int c, int waves, int tpf, const float* in,
                     float* out, int64_t nbatch, int use_weight, float scale, int cmode,
                     hipStream_t s, int layout) {
  const int n = p.class_n[c];
  if (n == 0 || nbatch <= 0) return HPX_OK;
  StageArgs a = make_args(p, nbatch, use_weight, scale, cmode, layout);
  a.plist = p.d_plist + p.class_off[c];
  if (c == kFftClasses - 1) {
    const int64_t items = (int64_t)n * nbatch;
    if (items > 0x7fffffff) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
    // every pair of the class fits the plan's longest FFT (at most kFftLds points)
    const size_t lds = (size_t)p.max_M * sizeof(float2);
    static const bool attr_set = [] {
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ring_analysis_block_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      (void)hipFuncSetAttribute(reinterpret_cast<const void*>(ring_synthesis_block_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
      return true;
    }();
    (void)attr_set;
    if (analysis)
      ring_analysis_block_kernel<<<(unsigned)items, kBlockThreads, lds, s>>>(in, out, a);
    else
      ring_synthesis_block_kernel<<<(unsigned)items, kBlockThreads, lds, s>>>(in, out, a);
    HPX_LAUNCH_CHECK();
    return HPX_OK;
  }
  // class c < 3: FFT length <= 2^(10 + c); tpf threads per field
  if (waves != kFftWaveChoices[c][0] && waves != kFftWaveChoices[c][1] && waves != kFftWaveChoices[c][2])
    HPX_FAIL(HPX_ERR_ARG, "no FFT kernel with %d fields per work-group for size class %d", waves, c);
  if ((tpf != 64 && tpf != 128 && tpf != 256) || waves * tpf > 1024)
    HPX_FAIL(HPX_ERR_ARG, "no FFT kernel with %d fields of %d threads", waves, tpf);
  const int64_t nfg = (nbatch + waves - 1) / waves;
  const int64_t items = (int64_t)n * nfg;
  if (items > 0x7fffffff) HPX_FAIL(HPX_ERR_UNSUPPORTED, "batch too large");
  const size_t lds_bytes = fft_lds_bytes(p, c, waves);
  if (lds_bytes > 160 * 1024) HPX_FAIL(HPX_ERR_UNSUPPORTED, "mmax too large for the ring FFT kernels");
  const unsigned grid = (unsigned)items;
#define HPX_W(W, C, T) \
  case (C * 32 + W) * 8 + T / 64: launch_wave<W, C, T>(analysis, grid, lds_bytes, in, out, a, s); break;
  switch ((c * 32 + waves) * 8 + tpf / 64) {
    HPX_W(4, 0, 64) HPX_W(8, 0, 64) HPX_W(16, 0, 64) HPX_W(4, 0, 128) HPX_W(8, 0, 128) HPX_W(4, 0, 256)
    HPX_W(2, 1, 64) HPX_W(4, 1, 64) HPX_W(8, 1, 64) HPX_W(2, 1, 128) HPX_W(4, 1, 128) HPX_W(8, 1, 128)
    HPX_W(2, 1, 256) HPX_W(4, 1, 256)
    HPX_W(1, 2, 64) HPX_W(2, 2, 64) HPX_W(4, 2, 64) HPX_W(1, 2, 128) HPX_W(2, 2, 128) HPX_W(4, 2, 128)
    HPX_W(1, 2, 256) HPX_W(2, 2, 256) HPX_W(4, 2, 256)
    default: HPX_FAIL(HPX_ERR_ARG, "no FFT kernel for class %d, %d fields, %d threads", c, waves, tpf);
  }
#undef HPX_W
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

namespace {

template <bool ANALYSIS>
int launch_stage(const ShtPlan& p, const float* in, float* out, int64_t nbatch, int use_weight,
                 float scale, int cmode, hipStream_t s, int layout) {
  if (nbatch <= 0) return HPX_OK;
  if (ANALYSIS && layout == 0 && p.kpad > p.npair) {
    zero_pad_rows_kernel<<<(unsigned)(2 * p.mmax), 256, 0, s>>>(out, p.kpad, p.npair, col_pad(nbatch));
    HPX_LAUNCH_CHECK();
  }
  const FftTune t = p.fft_tune(col_pad(nbatch));
  for (int c = 0; c < kFftClasses; ++c) {
    int waves = c < 3 ? t.waves[ANALYSIS ? 0 : 1][c] : 1;
    const int tpf = c < 3 ? t.tpf[ANALYSIS ? 0 : 1][c] : 64;
    // a large mmax (its phase factors share the work-group's LDS) leaves
    // room for fewer fields per work-group: take the largest that fits
    for (int j = 2; c < 3 && j > 0 && waves == kFftWaveChoices[c][j]; --j)
      if (fft_lds_bytes(p, c, waves) > 160 * 1024) waves = kFftWaveChoices[c][j - 1];
    const int rc = launch_fft_class(p, ANALYSIS, c, waves, tpf, in, out, nbatch, use_weight, scale, cmode, s,
                                    layout);
    if (rc != HPX_OK) return rc;
  }
  return HPX_OK;
}

}  // namespace

int fft_class_of(int log2M) {
  if (log2M <= 10) return 0;
  if (log2M == 11) return 1;
  if (log2M == 12) return 2;
  return 3;
}

int launch_analysis(const ShtPlan& p, const float* f, float* X, int64_t nbatch, int use_weight,
                    float scale, int cmode, hipStream_t s, int layout) {
  return launch_stage<true>(p, f, X, nbatch, use_weight, scale, cmode, s, layout);
}

int launch_synthesis(const ShtPlan& p, const float* X, float* f, int64_t nbatch, int use_weight,
                     float scale, int cmode, hipStream_t s, int layout) {
  return launch_stage<false>(p, X, f, nbatch, use_weight, scale, cmode, s, layout);
}

}  // namespace hpxmi
