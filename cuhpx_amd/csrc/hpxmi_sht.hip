// SHT plans and the C-ABI entry points.
//
// A plan holds what the reference rebuilds per call or per module:
//  * ring geometry of cuhpx/sht_tools.py:114-189, 319-332 (z, nphi, phi0,
//    offsets) and the ring quadrature weights (sht_tools.py:47-59, passed in);
//  * Bluestein tables for the rings whose length is not a power of two,
//    computed in FP64 on the host: chirp exp(-i pi (j^2 mod 2N) / N) (the
//    reduction mod 2N keeps the phase exact; the reference evaluates
//    float(PI)*j*j/N in FP32, hpx_fft_cuda.cu:1090) and the kernel spectrum
//    FFT_M(conj chirp) / M (the reference's y_pad FFT, hpx_fft.h:96-98);
//  * per ring-pair phase factors exp(-i m phi0) (hpx_fft_cuda.cu:166-179);
//  * the Legendre table, generated on the device (hpxmi_legendre.hip).

#include "hpxmi_sht_internal.h"
#include "../../include/hpxmi.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <mutex>
#include <string>
#include <vector>

namespace hpxmi {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

namespace {

using cd = std::complex<double>;

}  // namespace

// In-place iterative radix-2 FFT, exp(-2 pi i jk / n), n a power of two.
void host_fft(std::vector<std::complex<double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; ++i) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const size_t half = len / 2;
    std::vector<cd> w(half);
    for (size_t k = 0; k < half; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)len;
      w[k] = cd(std::cos(ang), std::sin(ang));
    }
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < half; ++k) {
        const cd u = a[i + k];
        const cd v = a[i + k + half] * w[k];
        a[i + k] = u + v;
        a[i + k + half] = u - v;
      }
  }
}

namespace {

bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

int ilog2(int64_t n) {
  int r = 0;
  while ((int64_t(1) << r) < n) ++r;
  return r;
}

float2 to_f2(cd z) { return make_float2((float)z.real(), (float)z.imag()); }

template <typename T>
int upload(T** dst, const std::vector<T>& src, int64_t& counter) {
  if (src.empty()) {
    *dst = nullptr;
    return HPX_OK;
  }
  const size_t bytes = src.size() * sizeof(T);
  HPX_HIP(hipMalloc(reinterpret_cast<void**>(dst), bytes));
  HPX_HIP(hipMemcpy(*dst, src.data(), bytes, hipMemcpyHostToDevice));
  counter += (int64_t)bytes;
  return HPX_OK;
}

void free_plan(ShtPlan* p) {
  if (p == nullptr) return;
  hipFree(p->d_pairs);
  hipFree(p->d_phase);
  hipFree(p->d_chirp);
  hipFree(p->d_bhat);
  hipFree(p->d_twiddle);
  hipFree(p->d_p);
  hipFree(p->d_pt);
  hipFree(p->d_p64);
  hipFree(p->d_pairs64);
  hipFree(p->d_chirp64);
  hipFree(p->d_bhat64);
  hipFree(p->d_tw64);
  hipFree(p->d_phase64);
  hipFree(p->d_items64);
  hipFree(p->d_w64);
  hipFree(p->d_ck);
  hipFree(p->d_ab);
  hipFree(p->d_seg_off);
  hipFree(p->d_ct);
  hipFree(p->d_p_off);
  hipFree(p->d_nl);
  hipFree(p->d_fwd_items);
  hipFree(p->d_kmin);
  hipFree(p->d_plist);
  delete p;
}

// Builds a plan from ring geometry; cos_theta[t] is used for the Legendre
// table of the north rings (t < npair).
// vector: build the two tables of the vector transform (D0, D1) instead of P.
int build_plan(int nring, const int32_t* nphi, const double* cos_theta, const double* phi0,
               const double* weights, int lmax, int mmax, hipStream_t stream, ShtPlan** out,
               int vector = 0, int csphase = 1, int vnorm = 0, int vinverse = 0) {
  if (nring < 1) HPX_FAIL(HPX_ERR_ARG, "nring must be >= 1");
  if (lmax < 0 || mmax < 1) HPX_FAIL(HPX_ERR_ARG, "lmax must be >= 0 and mmax >= 1 (got %d, %d)", lmax, mmax);
  if (mmax > 65535 || lmax > 65535) HPX_FAIL(HPX_ERR_UNSUPPORTED, "lmax/mmax too large");
  for (int t = 0; t < nring; ++t) {
    const int u = nring - 1 - t;
    if (nphi[t] < 1) HPX_FAIL(HPX_ERR_ARG, "ring %d has no pixels", t);
    if (nphi[t] != nphi[u]) HPX_FAIL(HPX_ERR_ARG, "grid is not mirror-symmetric (ring %d)", t);
    if (std::fabs(phi0[t] - phi0[u]) > 1e-12) HPX_FAIL(HPX_ERR_ARG, "ring phases are not mirror-symmetric");
    if (std::fabs(weights[t] - weights[u]) > 1e-12 * (std::fabs(weights[t]) + 1e-300))
      HPX_FAIL(HPX_ERR_ARG, "ring weights are not mirror-symmetric");
    if (std::fabs(cos_theta[t] + cos_theta[u]) > 1e-9) HPX_FAIL(HPX_ERR_ARG, "ring latitudes are not mirror-symmetric");
  }

  ShtPlan* p = new ShtPlan();
  p->nring = nring;
  p->npair = (nring + 1) / 2;
  p->kpad = (p->npair + kKTile - 1) / kKTile * kKTile;
  p->lmax = lmax;
  p->mmax = mmax;
  p->vector = vector ? 1 : 0;
  p->csphase = csphase ? 1 : 0;
  p->vnorm = vnorm;
  p->vinverse = vinverse ? 1 : 0;

  std::vector<int64_t> offset(nring + 1, 0);
  for (int t = 0; t < nring; ++t) offset[t + 1] = offset[t] + nphi[t];
  p->npix = offset[nring];

  std::vector<float2> chirp, bhat;
  p->pairs.resize(p->npair);
  int max_M = 1;
  for (int k = 0; k < p->npair; ++k) {
    RingPair& r = p->pairs[k];
    const int N = nphi[k];
    r.nphi = N;
    // FFT length for a sequence of n points: n itself when a power of two,
    // else the Bluestein length
    auto fft_len = [](int64_t n) { return is_pow2(n) ? n : int64_t(1) << ilog2(2 * n - 1); };
    r.M = (int)std::min<int64_t>(fft_len(N), int64_t(1) << 30);
    // A ring whose pair transform would need a longer FFT than LDS holds is
    // transformed ring by ring as N/2 complex points z_j = f_2j + i f_2j+1
    // (half = 1), and if that is still too long, as the two interleaved
    // halves of z, N/4 points each (half = 2, Bluestein only).
    r.half = 0;
    if (r.M > kFftLds && N % 2 == 0 && fft_len(N / 2) <= kFftLds) {
      r.half = 1;
      r.M = (int)fft_len(N / 2);
    } else if (r.M > kFftLds && N % 4 == 0 && !is_pow2(N / 4) && fft_len(N / 4) <= kFftLds) {
      r.half = 2;
      r.M = (int)fft_len(N / 4);
    }
    r.nseq = N >> r.half;
    if (r.M > kFftLds) {
      free_plan(p);
      HPX_FAIL(HPX_ERR_UNSUPPORTED, "ring of %d pixels needs an FFT of %d > %d points", N, r.M, kFftLds);
    }
    r.log2M = ilog2(r.M);
    r.single = (k == nring - 1 - k) ? 1 : 0;
    r.off_n = offset[k];
    r.off_s = offset[nring - 1 - k];
    r.weight = (float)weights[k];
    r.mwrite = mmax;
    r.chirp_off = 0;
    r.bhat_off = 0;
    max_M = std::max(max_M, r.M);
    if (r.M != r.nseq) {
      r.chirp_off = (int)chirp.size();
      r.bhat_off = (int)bhat.size();
      const int Nc = r.nseq;  // length of the transformed sequence
      std::vector<cd> w(Nc);
      for (int j = 0; j < Nc; ++j) {
        const int64_t q = ((int64_t)j * j) % (2 * (int64_t)Nc);
        const double ang = -M_PI * (double)q / (double)Nc;
        w[j] = cd(std::cos(ang), std::sin(ang));
        chirp.push_back(to_f2(w[j]));
      }
      std::vector<cd> b(r.M, cd(0.0, 0.0));
      b[0] = std::conj(w[0]);
      for (int j = 1; j < Nc; ++j) {
        b[j] = std::conj(w[j]);
        b[r.M - j] = std::conj(w[j]);
      }
      host_fft(b);
      for (int x = 0; x < r.M; ++x) bhat.push_back(to_f2(b[x] / (double)r.M));
    }
  }
  p->max_M = max_M;
  p->log2_max_M = ilog2(max_M);

  // Ring pairs grouped by FFT size class, in ring order.  (A cost-balanced
  // order, pairs sorted by estimated FFT cost and dealt round-robin over the
  // 8 XCDs' contiguous ranges, sped up the 4- and 8-field work-groups but
  // slowed the 16-field ones that the tuner picks at nside 256.)
  std::vector<int> plist;
  for (int c = 0; c < kFftClasses; ++c) {
    p->class_off[c] = (int)plist.size();
    for (int k = 0; k < p->npair; ++k)
      if (fft_class_of(p->pairs[k].log2M) == c) plist.push_back(k);
    p->class_n[c] = (int)plist.size() - p->class_off[c];
  }

  std::vector<float2> tw(max_M);
  for (int x = 0; x < max_M; ++x) {
    const double ang = -2.0 * M_PI * (double)x / (double)max_M;
    tw[x] = make_float2((float)std::cos(ang), (float)std::sin(ang));
  }

  std::vector<float2> phase((size_t)p->npair * mmax);
  for (int k = 0; k < p->npair; ++k)
    for (int m = 0; m < mmax; ++m) {
      const double ang = std::fmod((double)m * phi0[k], 2.0 * M_PI);
      phase[(size_t)k * mmax + m] = make_float2((float)std::cos(ang), (float)-std::sin(ang));
    }

  // Legendre table layout
  p->nl.resize(2 * (size_t)mmax);
  p->p_off.resize(mmax + 1);
  p->p_off[0] = 0;
  std::vector<int2> items;
  int max_blocks = 0;
  for (int m = 0; m < mmax; ++m) {
    for (int par = 0; par < 2; ++par) {
      const int n = (lmax - m - par + 1) / 2;
      p->nl[2 * m + par] = n > 0 ? n : 0;
    }
    const int pad0 = (p->nl[2 * m] + kRowPad - 1) / kRowPad * kRowPad;
    const int pad1 = (p->nl[2 * m + 1] + kRowPad - 1) / kRowPad * kRowPad;
    p->p_off[m + 1] = p->p_off[m] + (int64_t)(pad0 + pad1) * p->kpad;
    const int nmax = std::max(p->nl[2 * m], p->nl[2 * m + 1]);
    max_blocks = std::max(max_blocks, (nmax + 63) / 64);
  }
  // forward work list: m outer, row block inner.  The work-groups of one m
  // read the same spectra X[m] (512 KB at 64 fields); the launch maps
  // contiguous ranges of this list to one XCD, so they share that XCD's L2.
  // Both the row-block count and (with the polar ring pairs skipped, see
  // kmin below) the reduction length fall with m, so m runs in the order 0,
  // mmax-1, 1, mmax-2, ...: every contiguous eighth then carries about the
  // same work.
  for (int mi = 0; mi < mmax; ++mi) {
    const int m = (mi & 1) ? mmax - 1 - (mi >> 1) : (mi >> 1);
    const int nmax = std::max(p->nl[2 * m], p->nl[2 * m + 1]);
    for (int jb = 0; jb * 64 < nmax; ++jb) items.push_back(make_int2(m, jb));
  }
  (void)max_blocks;
  p->n_fwd_items = (int)items.size();
  // kmin[m]: the first ring pair k with K_k = min(mmax, nphi_k / 2 + 1) > m.
  // The analysis writes zero spectra for m >= K_k and the synthesis reads
  // only m < K_k, so the forward GEMM skips the k chunks wholly below
  // kmin[m] and the inverse GEMM the ring-pair blocks wholly below it (the
  // polar rings at high m: a quarter of the spectra at nside 256, L 513).
  std::vector<int> kmin(mmax, p->npair);
  for (int k = p->npair - 1; k >= 0; --k) {
    const int K = std::min(mmax, p->pairs[k].nphi / 2 + 1);
    for (int m = 0; m < K; ++m) kmin[m] = k;
  }
  // The forward GEMM reads ring pair k at order m when k's chunk of kKTile
  // pairs is not below kmin[m]'s; the analysis writes pair k only up to the
  // last such m (its zeros beyond K_k included), no further.
  {
    std::vector<int> mread(p->npair, 0);
    for (int m = 0; m < mmax; ++m)
      for (int k = kmin[m] / kKTile * kKTile; k < p->npair; ++k) mread[k] = std::max(mread[k], m + 1);
    for (int k = 0; k < p->npair; ++k) {
      const int K = std::min(mmax, p->pairs[k].nphi / 2 + 1);
      p->pairs[k].mwrite = std::min(mmax, std::max(K, mread[k]));
    }
  }

  std::vector<double> ct(p->kpad, 0.0);
  for (int k = 0; k < p->npair; ++k) ct[k] = cos_theta[k];
  p->h_ct.assign(cos_theta, cos_theta + p->npair);
  p->h_phi0.assign(phi0, phi0 + p->npair);
  p->h_w.assign(weights, weights + p->npair);
  // Scalar plans: recurrence segments of 16 degrees per order, and the
  // coefficients (a_l, b_l) of P_l = a_l x P_{l-1} - b_l P_{l-2}
  // (sht_tools.py:250-251) per segment degree, FP64-computed, zero for the
  // segment's first two degrees (stored values) and from lmax on.
  std::vector<float2> ab;
  std::vector<float> ctf;
  if (!p->vector) {
    p->seg_off.assign(mmax + 1, 0);
    for (int m = 0; m < mmax; ++m) p->seg_off[m + 1] = p->seg_off[m] + nseg_of(lmax, m);
    ab.assign((size_t)p->seg_off[mmax] * kSegL, make_float2(0.f, 0.f));
    for (int m = 0; m < mmax; ++m)
      for (int t = 0; t < nseg_of(lmax, m); ++t)
        for (int j = 2; j < kSegL; ++j) {
          const int l = m + kSegL * t + j;
          if (l >= lmax) break;
          const double dl = l, dm = m;
          const double a = std::sqrt((2 * dl - 1) / (dl - dm) * (2 * dl + 1) / (dl + dm));
          const double b = std::sqrt((dl + dm - 1) / (dl - dm) * (2 * dl + 1) / (2 * dl - 3) * (dl - dm - 1) / (dl + dm));
          ab[((size_t)p->seg_off[m] + t) * kSegL + j] = make_float2((float)a, (float)b);
        }
    ctf.assign(p->kpad, 0.f);
    for (int k = 0; k < p->npair; ++k) ctf[k] = (float)cos_theta[k];
  }

  int64_t bytes = 0;
  int rc;
  if ((rc = upload(&p->d_pairs, p->pairs, bytes)) != HPX_OK ||
      (rc = upload(&p->d_phase, phase, bytes)) != HPX_OK ||
      (rc = upload(&p->d_chirp, chirp, bytes)) != HPX_OK ||
      (rc = upload(&p->d_bhat, bhat, bytes)) != HPX_OK ||
      (rc = upload(&p->d_twiddle, tw, bytes)) != HPX_OK ||
      (rc = upload(&p->d_p_off, p->p_off, bytes)) != HPX_OK ||
      (rc = upload(&p->d_nl, p->nl, bytes)) != HPX_OK ||
      (rc = upload(&p->d_fwd_items, items, bytes)) != HPX_OK ||
      (rc = upload(&p->d_kmin, kmin, bytes)) != HPX_OK ||
      (rc = upload(&p->d_plist, plist, bytes)) != HPX_OK ||
      (rc = upload(&p->d_seg_off, p->seg_off, bytes)) != HPX_OK ||
      (rc = upload(&p->d_ab, ab, bytes)) != HPX_OK ||
      (rc = upload(&p->d_ct, ctf, bytes)) != HPX_OK) {
    free_plan(p);
    return rc;
  }
  // the FFT kernels read chirp / bhat through pointers even when unused
  if (p->d_chirp == nullptr) {
    std::vector<float2> one(1, make_float2(1.f, 0.f));
    if ((rc = upload(&p->d_chirp, one, bytes)) != HPX_OK || (rc = upload(&p->d_bhat, one, bytes)) != HPX_OK) {
      free_plan(p);
      return rc;
    }
  }
  const int64_t pfloats = p->p_off[mmax];
  p->at_rows = pfloats / p->kpad;
  p->vstride = p->vector ? pfloats : 0;
  // vector plans: the two tables and their transposes; scalar plans: the
  // segment checkpoints
  const int64_t tfloats = std::max<int64_t>((p->vector ? 2 : 1) * pfloats, 1);
  const int64_t big_bytes = p->vector ? 2 * tfloats * (int64_t)sizeof(float)
                                      : std::max<int64_t>((int64_t)p->seg_off[mmax] * p->kpad, 1) * (int64_t)sizeof(float2);
  {
    // the plan's only large allocation: refuse it with the sizes named
    // rather than fail inside hipMalloc
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && (uint64_t)big_bytes > (uint64_t)free_b) {
      free_plan(p);
      HPX_FAIL(HPX_ERR_UNSUPPORTED, "plan tables need %.2f GB, %.2f GB free of %.2f GB on the device",
               big_bytes / 1e9, free_b / 1e9, total_b / 1e9);
    }
  }
  double* d_ct = nullptr;
  hipError_t e = hipSuccess;
  if (p->vector) {
    e = hipMalloc(reinterpret_cast<void**>(&p->d_p), tfloats * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&p->d_pt), tfloats * sizeof(float));
  } else {
    e = hipMalloc(reinterpret_cast<void**>(&p->d_ck), big_bytes);
  }
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&d_ct), ct.size() * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(d_ct, ct.data(), ct.size() * sizeof(double), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(d_ct);
    free_plan(p);
    HPX_FAIL(HPX_ERR_HIP, "plan allocation failed: %s", hipGetErrorString(e));
  }
  bytes += big_bytes;
  rc = launch_legendre_table(*p, d_ct, stream);
  if (rc == HPX_OK) {
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) {
      set_error(std::string("Legendre table generation failed: ") + hipGetErrorString(e));
      rc = HPX_ERR_HIP;
    }
  }
  hipFree(d_ct);
  if (rc != HPX_OK) {
    free_plan(p);
    return rc;
  }
  p->table_bytes = bytes;
  *out = p;
  return HPX_OK;
}

// nbatch: scalar fields (2 per vector field for the vector calls)
int check_call(const ShtPlan* p, int precision, int64_t nbatch, int64_t ws_bytes, bool vector_call = false) {
  if (p == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (p->lmax < 1) HPX_FAIL(HPX_ERR_ARG, "plan was built with lmax = 0 (ring FFTs only)");
  if (p->vector && !vector_call) HPX_FAIL(HPX_ERR_ARG, "plan holds vector-transform tables (use hpx_vsht_*)");
  if (!p->vector && vector_call) HPX_FAIL(HPX_ERR_ARG, "plan holds the scalar table (use hpx_vsht_plan_create_healpix)");
  if (precision < 0 || precision > 2) HPX_FAIL(HPX_ERR_ARG, "unknown precision %d", precision);
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch > max_maps(*p))
    HPX_FAIL(HPX_ERR_UNSUPPORTED, "%lld maps in one call exceed this plan's limit of %lld; split the batch",
             (long long)nbatch, (long long)max_maps(*p));
  const int64_t need = workspace_floats(*p, nbatch) * (int64_t)sizeof(float);
  if (ws_bytes < need)
    HPX_FAIL(HPX_ERR_ARG, "workspace of %lld bytes is smaller than the %lld bytes needed",
             (long long)ws_bytes, (long long)need);
  return HPX_OK;
}

}  // namespace
}  // namespace hpxmi

using namespace hpxmi;

namespace {
ShtPlan* P(const hpx_sht_plan* h) { return reinterpret_cast<ShtPlan*>(const_cast<hpx_sht_plan*>(h)); }
hpx_sht_plan* H(ShtPlan* p) { return reinterpret_cast<hpx_sht_plan*>(p); }
}  // namespace

extern "C" const char* hpx_last_error(void) { return hpxmi::last_error(); }

extern "C" int hpx_version(void) { return 100; }

extern "C" int hpx_sht_plan_create_rings(int nring, const int32_t* nphi, const double* theta,
                                         const double* phi0, const double* weights, int lmax,
                                         int mmax, hipStream_t stream, hpx_sht_plan** plan) {
  if (plan == nullptr || nphi == nullptr || theta == nullptr || phi0 == nullptr || weights == nullptr)
    HPX_FAIL(HPX_ERR_ARG, "null argument");
  std::vector<double> ct(nring > 0 ? nring : 0);
  for (int t = 0; t < nring; ++t) ct[t] = std::cos(theta[t]);
  ShtPlan* p = nullptr;
  const int rc = build_plan(nring, nphi, ct.data(), phi0, weights, lmax, mmax, stream, &p);
  if (rc != HPX_OK) return rc;
  *plan = H(p);
  return HPX_OK;
}

namespace {
int healpix_plan(int nside, int lmax, int mmax, const double* ring_weights, hipStream_t stream,
                 hpx_sht_plan** plan, int vector, int csphase, int vnorm = 0, int vinverse = 0) {
  if (plan == nullptr || ring_weights == nullptr) HPX_FAIL(HPX_ERR_ARG, "null argument");
  if (nside < 1) HPX_FAIL(HPX_ERR_ARG, "nside must be >= 1");
  if (nside > 8192) HPX_FAIL(HPX_ERR_UNSUPPORTED, "nside > 8192 exceeds the FFT size limit");
  const int nring = 4 * nside - 1;
  std::vector<int32_t> nphi(nring);
  std::vector<double> z(nring), phi0(nring);
  const double ns = nside;
  for (int t = 0; t < nring; ++t) {
    // sht_tools.py:179-189 (nphi), 129-136 (z, before the flip), 161-176 (phi0)
    if (t < nside - 1) {
      nphi[t] = 4 * (t + 1);
      z[t] = 1.0 - (double)(t + 1) * (t + 1) / (3.0 * ns * ns);
    } else if (t <= 3 * nside - 1) {
      nphi[t] = 4 * nside;
      z[t] = 4.0 / 3.0 - 2.0 * (t + 1) / (3.0 * ns);
    } else {
      nphi[t] = 4 * (4 * nside - t - 1);
      z[t] = (double)(4 * nside - 1 - t) * (4 * nside - 1 - t) / (3.0 * ns * ns) - 1.0;
    }
    if (t + 1 >= nside && t + 1 <= 3 * nside)
      phi0[t] = M_PI / (2.0 * ns) * (0.5 * ((t - nside + 2) % 2));
    else if (t + 1 > 3 * nside)
      phi0[t] = M_PI / (2.0 * (4 * nside - t - 1)) * 0.5;
    else
      phi0[t] = M_PI / (2.0 * (t + 1)) * 0.5;
  }
  ShtPlan* p = nullptr;
  const int rc = build_plan(nring, nphi.data(), z.data(), phi0.data(), ring_weights, lmax, mmax, stream, &p,
                            vector, csphase, vnorm, vinverse);
  if (rc != HPX_OK) return rc;
  *plan = H(p);
  return HPX_OK;
}
}  // namespace

extern "C" int hpx_sht_plan_create_healpix(int nside, int lmax, int mmax, const double* ring_weights,
                                           hipStream_t stream, hpx_sht_plan** plan) {
  return healpix_plan(nside, lmax, mmax, ring_weights, stream, plan, 0, 1);
}

extern "C" int hpx_vsht_plan_create_healpix(int nside, int lmax, int mmax, const double* ring_weights,
                                            int csphase, int norm, int inverse, hipStream_t stream,
                                            hpx_sht_plan** plan) {
  if (lmax < 1) HPX_FAIL(HPX_ERR_ARG, "a vector plan needs lmax >= 1");
  if (norm != HPX_NORM_ORTHO && norm != HPX_NORM_FOUR_PI && norm != HPX_NORM_SCHMIDT)
    HPX_FAIL(HPX_ERR_ARG, "unknown norm %d", norm);
  return healpix_plan(nside, lmax, mmax, ring_weights, stream, plan, 1, csphase, norm,
                      norm == HPX_NORM_ORTHO ? 0 : inverse);
}

extern "C" int64_t hpx_vsht_workspace_bytes(const hpx_sht_plan* plan, int64_t nbatch) {
  if (plan == nullptr || nbatch < 0 || nbatch > (INT64_MAX >> 2)) return -1;
  return workspace_floats(*P(plan), 2 * nbatch) * (int64_t)sizeof(float);
}

extern "C" int hpx_vsht_forward(const hpx_sht_plan* plan, int precision, const float* f, float* alm,
                                int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  if (nbatch < 0 || nbatch > (INT64_MAX >> 2)) HPX_FAIL(HPX_ERR_ARG, "bad nbatch");
  int rc = check_call(pl, precision, 2 * nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  // the two components of each field are consecutive maps: 2 nbatch scalar
  // ring analyses (quadrature weight applied), then the vector GEMM
  float* X = static_cast<float*>(workspace);
  if ((rc = launch_analysis(*pl, f, X, 2 * nbatch, 1, 1.0f, 0, stream)) != HPX_OK) return rc;
  return launch_legendre_fwd(*pl, precision, X, alm, 2 * nbatch, stream);
}

extern "C" int hpx_vsht_inverse(const hpx_sht_plan* plan, int precision, const float* alm, float* f,
                                int64_t nbatch, void* workspace, int64_t workspace_bytes,
                                hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  if (nbatch < 0 || nbatch > (INT64_MAX >> 2)) HPX_FAIL(HPX_ERR_ARG, "bad nbatch");
  int rc = check_call(pl, precision, 2 * nbatch, workspace_bytes, true);
  if (rc != HPX_OK || nbatch == 0) return rc;
  float* X = static_cast<float*>(workspace);
  if ((rc = launch_legendre_inv(*pl, precision, alm, X, 2 * nbatch, stream)) != HPX_OK) return rc;
  return launch_synthesis(*pl, X, f, 2 * nbatch, 0, 1.0f, 0, stream);
}

extern "C" void hpx_sht_plan_destroy(hpx_sht_plan* plan) { free_plan(P(plan)); }

extern "C" int64_t hpx_sht_plan_npix(const hpx_sht_plan* plan) { return plan ? P(plan)->npix : -1; }

extern "C" int64_t hpx_sht_plan_table_bytes(const hpx_sht_plan* plan) {
  return plan ? P(plan)->table_bytes : -1;
}

extern "C" int64_t hpx_sht_workspace_bytes(const hpx_sht_plan* plan, int64_t nbatch) {
  if (plan == nullptr || nbatch < 0) return -1;
  return workspace_floats(*P(plan), nbatch) * (int64_t)sizeof(float);
}

extern "C" int64_t hpx_sht_max_batch(const hpx_sht_plan* plan) {
  if (plan == nullptr) return -1;
  const ShtPlan* p = P(plan);
  return p->vector ? max_maps(*p) / 2 : max_maps(*p);
}

extern "C" int hpx_sht_forward(const hpx_sht_plan* plan, int precision, const float* f, float* alm,
                               int64_t nbatch, void* workspace, int64_t workspace_bytes,
                               hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  float* X = static_cast<float*>(workspace);
  if ((rc = launch_analysis(*pl, f, X, nbatch, 1, 1.0f, 0, stream)) != HPX_OK) return rc;
  return launch_legendre_fwd(*pl, precision, X, alm, nbatch, stream);
}

// ---------------------------------------------------------------------------
// Forward SHT without a caller workspace: the workspace is a fixed device
// array of the library (its address looked up once per device), for callers
// that cannot allocate device memory.
// ---------------------------------------------------------------------------

static constexpr int64_t kScratchBytes = int64_t(128) << 20;
// Held by every device the library runs on until the process ends.  Host
// code never uses this name as an address: scratch_ptr() gives the device's
// address of its copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device ordinal

// The current device's address of `scratch`, looked up on the first call on
// that device.
int scratch_ptr(void** out) {
  int dev = 0;
  HPX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (dev >= (int)g_scratch_ptr.size()) g_scratch_ptr.resize(dev + 1, nullptr);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    HPX_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (ptr == nullptr) HPX_FAIL(HPX_ERR_HIP, "no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return HPX_OK;
}
}  // namespace

extern "C" int hpx_sht_forward_noalloc(const hpx_sht_plan* plan, int precision, const float* f, float* alm,
                                       int64_t nbatch, hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  if (pl == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch > max_maps(*pl))
    HPX_FAIL(HPX_ERR_UNSUPPORTED, "%lld maps in one call exceed this plan's limit of %lld; split the batch",
             (long long)nbatch, (long long)max_maps(*pl));
  const int64_t need = workspace_floats(*pl, nbatch) * (int64_t)sizeof(float);
  if (need > kScratchBytes)
    HPX_FAIL(HPX_ERR_UNSUPPORTED,
             "%lld maps need %lld bytes of workspace, more than the %lld-byte scratch of the "
             "allocation-free entry point; split the batch or use hpx_sht_forward",
             (long long)nbatch, (long long)need, (long long)kScratchBytes);
  void* ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc != HPX_OK) return rc;
  // every stage reads only workspace entries an earlier stage of the same
  // call wrote, so nothing needs clearing between calls
  return hpx_sht_forward(plan, precision, f, alm, nbatch, ws, kScratchBytes, stream);
}

extern "C" int hpx_sht_inverse(const hpx_sht_plan* plan, int precision, const float* alm, float* f,
                               int64_t nbatch, void* workspace, int64_t workspace_bytes,
                               hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  float* X = static_cast<float*>(workspace);
  if ((rc = launch_legendre_inv(*pl, precision, alm, X, nbatch, stream)) != HPX_OK) return rc;
  return launch_synthesis(*pl, X, f, nbatch, 0, 1.0f, 0, stream);
}

extern "C" int hpx_sht_forward_adjoint(const hpx_sht_plan* plan, int precision, int mode,
                                       const float* galm, float* gf, int64_t nbatch,
                                       void* workspace, int64_t workspace_bytes,
                                       hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  if (mode != HPX_ADJ_EXACT && mode != HPX_ADJ_COMPAT) HPX_FAIL(HPX_ERR_ARG, "unknown adjoint mode %d", mode);
  float* X = static_cast<float*>(workspace);
  if ((rc = launch_legendre_inv(*pl, precision, galm, X, nbatch, stream)) != HPX_OK) return rc;
  // exact: w_t * sum_{m<K} Re(H e^{i m phi});  compat: (w_t / 2) * irfft(H)
  if (mode == HPX_ADJ_EXACT) return launch_synthesis(*pl, X, gf, nbatch, 1, 1.0f, 1, stream);
  return launch_synthesis(*pl, X, gf, nbatch, 1, 0.5f, 0, stream);
}

extern "C" int hpx_sht_inverse_adjoint(const hpx_sht_plan* plan, int precision, int mode,
                                       const float* gf, float* galm, int64_t nbatch,
                                       void* workspace, int64_t workspace_bytes,
                                       hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  if (mode != HPX_ADJ_EXACT && mode != HPX_ADJ_COMPAT) HPX_FAIL(HPX_ERR_ARG, "unknown adjoint mode %d", mode);
  float* X = static_cast<float*>(workspace);
  // exact: c_m * rfft(g);  compat: 2 * rfft(g)
  if (mode == HPX_ADJ_EXACT) rc = launch_analysis(*pl, gf, X, nbatch, 0, 1.0f, 1, stream);
  else rc = launch_analysis(*pl, gf, X, nbatch, 0, 2.0f, 0, stream);
  if (rc != HPX_OK) return rc;
  return launch_legendre_fwd(*pl, precision, X, galm, nbatch, stream);
}

extern "C" int hpx_sht_stage_ring_analysis(const hpx_sht_plan* plan, const float* f, int64_t nbatch,
                                           void* workspace, int64_t workspace_bytes,
                                           hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, 0, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_analysis(*pl, f, static_cast<float*>(workspace), nbatch, 1, 1.0f, 0, stream);
}

extern "C" int hpx_sht_stage_legendre_analysis(const hpx_sht_plan* plan, int precision, float* alm,
                                               int64_t nbatch, void* workspace,
                                               int64_t workspace_bytes, hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_legendre_fwd(*pl, precision, static_cast<float*>(workspace), alm, nbatch, stream);
}

extern "C" int hpx_sht_stage_legendre_synthesis(const hpx_sht_plan* plan, int precision,
                                                const float* alm, int64_t nbatch, void* workspace,
                                                int64_t workspace_bytes, hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, precision, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_legendre_inv(*pl, precision, alm, static_cast<float*>(workspace), nbatch, stream);
}

extern "C" int hpx_sht_stage_ring_synthesis(const hpx_sht_plan* plan, float* f, int64_t nbatch,
                                            void* workspace, int64_t workspace_bytes,
                                            hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  int rc = check_call(pl, 0, nbatch, workspace_bytes);
  if (rc != HPX_OK || nbatch == 0) return rc;
  return launch_synthesis(*pl, static_cast<const float*>(workspace), f, nbatch, 0, 1.0f, 0, stream);
}

extern "C" int hpx_ring_rfft(const hpx_sht_plan* plan, const float* f, float* ftm, int64_t nbatch,
                             hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  if (pl == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch == 0) return HPX_OK;
  return launch_analysis(*pl, f, ftm, nbatch, 0, 1.0f, 0, stream, 1);
}

extern "C" int hpx_ring_irfft(const hpx_sht_plan* plan, const float* ftm, float* f, int64_t nbatch,
                              hipStream_t stream) {
  const ShtPlan* pl = P(plan);
  if (pl == nullptr) HPX_FAIL(HPX_ERR_ARG, "null plan");
  if (nbatch < 0) HPX_FAIL(HPX_ERR_ARG, "nbatch must be >= 0");
  if (nbatch == 0) return HPX_OK;
  return launch_synthesis(*pl, ftm, f, nbatch, 0, 1.0f, 0, stream, 1);
}
