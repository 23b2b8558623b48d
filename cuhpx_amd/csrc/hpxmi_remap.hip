// RING / NESTED / flat pixel remaps for gfx950.
//
// Every remap kind is a permutation of the 12*nside^2 pixels of a map that
// is the same for every row of the batch.  The index arithmetic follows the
// reference's kernels (src/data_remapping/hpx_remapping_cuda.cu):
//   ring <-> (ix, iy, face)  161-282,  nest <-> (ix, iy, face)  285-347,
//   flat <-> (ix, iy, face) with correct_xy_orient  33-134,
// and the scatter semantics out[sigma(i)] = in[i] (137-146, 902-917).
//
// MI355X design: instead of materialising a pix array and scattering every
// row through it (4 cudaMalloc'd index arrays, 4 launches, uncoalesced writes
// in the reference), one launch walks the map in tiles.  A tile is a
// TB x TB block of one base face (TB = min(nside, 64 or 128)), which is a contiguous
// Morton range in NEST order, TB rows of TB in the flat layout, and 2*TB-1
// contiguous ring segments in RING order.  Each work-group computes the
// source and destination index of its tile's pixels once, in registers, and
// then streams the batch rows through LDS: the source side is read in its
// own natural order (coalesced), the destination side is written in its own
// natural order (coalesced).  From nside 64 (RING kinds: 128) the tiles are
// 64 x 64 and each work-group's waves split by role (remap_tiles_split): half
// of them load row r into one LDS buffer while the other half stores row
// r - 1 from the other.  NEST -> RING at nside >= 512, and flat -> RING with
// 4-byte pixels there, use tiles that follow the rings instead
// (nest2ring_bands), so that their RING writes come in runs of 1-2 KB.

#include "hpxmi_common.h"

#include <algorithm>
#include "../../include/hpxmi.h"

namespace {


enum Scheme { kNest = 0, kRing = 1, kFlat = 2 };

struct Side {
  int scheme;
  int flip;  // flat only: swap ix <-> iy before rotating
  int k;     // flat only: quarter turns
};

__device__ __forceinline__ int spread_bits(int v) {
  int r = v & 0xffff;
  r = (r ^ (r << 8)) & 0x00ff00ff;
  r = (r ^ (r << 4)) & 0x0f0f0f0f;
  r = (r ^ (r << 2)) & 0x33333333;
  r = (r ^ (r << 1)) & 0x55555555;
  return r;
}

__device__ __forceinline__ int compress_bits(int v) {
  int r = v & 0x55555555;
  r = (r ^ (r >> 1)) & 0x33333333;
  r = (r ^ (r >> 2)) & 0x0f0f0f0f;
  r = (r ^ (r >> 4)) & 0x00ff00ff;
  r = (r ^ (r >> 8)) & 0x0000ffff;
  return r;
}

// One quarter turn of correct_xy_orient (k == 1), masked to the face.
__host__ __device__ __forceinline__ void rot1(int& x, int& y, int mask) {
  int nx = (-y - 1) & mask;
  int ny = x & mask;
  x = nx;
  y = ny;
}

// correct_xy_orient: optional swap, then k quarter turns.
__host__ __device__ __forceinline__ void correct(int& x, int& y, int flip, int k, int mask) {
  if (flip) {
    int t = x;
    x = y;
    y = t;
  }
  for (int i = 0; i < k; ++i) rot1(x, y, mask);
}

// Inverse of correct(): (4 - k) quarter turns, then the optional swap.
__host__ __device__ __forceinline__ void uncorrect(int& x, int& y, int flip, int k, int mask) {
  for (int i = 0; i < ((4 - k) & 3); ++i) rot1(x, y, mask);
  if (flip) {
    int t = x;
    x = y;
    y = t;
  }
}

// (ix, iy, face) -> RING index, as xyf2ring_kernel (hpx_remapping_cuda.cu:228-282).
__device__ __forceinline__ int64_t xyf2ring(int ix, int iy, int face, int64_t nside) {
  const int64_t nl4 = 4 * nside;
  const int64_t npface = nside * nside;
  const int64_t ncap = (npface - nside) << 1;
  const int64_t npix = 12 * npface;
  const int64_t jr = ((face >> 2) + 2) * nside - ix - iy - 1;
  int64_t nr, n_before;
  bool shifted;
  if (jr < nside) {
    shifted = true;
    nr = 4 * jr;
    n_before = 2 * jr * (jr - 1);
  } else if (jr < 3 * nside) {
    shifted = ((jr - nside) & 1) == 0;
    nr = 4 * nside;
    n_before = ncap + (jr - nside) * nr;
  } else {
    shifted = true;
    const int64_t ring_nr = 4 * nside - jr;
    nr = 4 * ring_nr;
    n_before = npix - 2 * ring_nr * (ring_nr + 1);
  }
  nr >>= 2;
  const int64_t kshift = shifted ? 0 : 1;
  int64_t jp = ((face & 3) << 1) - ((face >> 2) & 1) + 1;
  jp = (jp * nr + ix - iy + 1 + kshift) / 2;  // C division, as the reference
  if (jp < 1) jp += nl4;
  return n_before + jp - 1;
}

struct TileGeom {
  int face;
  int x0, y0;         // tile corner in face coordinates
  int64_t nest_base;  // NEST index of the tile's first pixel
  int tb, log_tb;
};

// Tile-local coordinates (lx, ly) and global index of the e-th pixel of the
// tile in the natural order of `side`.
__device__ __forceinline__ void element(const Side& side, const TileGeom& g, int e,
                                        int64_t nside, int order, int fx0, int fy0,
                                        int& lx, int& ly, int64_t& gidx) {
  const int tb = g.tb;
  const int mask = (int)nside - 1;
  if (side.scheme == kNest) {
    lx = compress_bits(e);
    ly = compress_bits(e >> 1);
    gidx = g.nest_base + e;
  } else if (side.scheme == kFlat) {
    // Row-major over the tile's image block in the flat layout.
    const int r = e >> g.log_tb;
    const int c = e & (tb - 1);
    int x = fx0 + c;
    int y = fy0 + r;
    gidx = ((int64_t)g.face << (2 * order)) + ((int64_t)y << order) + x;
    uncorrect(x, y, side.flip, side.k, mask);
    lx = x - g.x0;
    ly = y - g.y0;
  } else {
    // Ring segments s = lx + ly, each ordered by increasing lx (increasing
    // position along the ring); first the lower triangle, then the upper.
    const int tri = tb * (tb + 1) / 2;
    if (e < tri) {
      int s = (int)((sqrtf(8.0f * e + 1.0f) - 1.0f) * 0.5f);
      while ((s + 1) * (s + 2) / 2 <= e) ++s;
      while (s * (s + 1) / 2 > e) --s;
      const int pos = e - s * (s + 1) / 2;
      lx = pos;
      ly = s - pos;
    } else {
      const int ep = tb * tb - 1 - e;
      int s = (int)((sqrtf(8.0f * ep + 1.0f) - 1.0f) * 0.5f);
      while ((s + 1) * (s + 2) / 2 <= ep) ++s;
      while (s * (s + 1) / 2 > ep) --s;
      const int pos = ep - s * (s + 1) / 2;
      lx = tb - 1 - pos;
      ly = tb - 1 - (s - pos);
    }
    gidx = xyf2ring(g.x0 + lx, g.y0 + ly, g.face, nside);
  }
}

// Maps below nside 64 (RING kinds: 128), where the tile is the whole face
// side: TBMAX is the largest tile side (the work-group has TBMAX^2 / PT
// threads, PT pixels each, both loading and storing).  NTL / NTS:
// non-temporal loads / stores, for a side that streams whole lines of the
// tile (NEST or flat); the partial-line runs of a RING side go through the
// L2, which merges them.
template <typename T, int TBMAX, int PT, bool NTL, bool NTS>
__global__ __launch_bounds__(TBMAX * TBMAX / PT) void remap_tiles(
    const T* __restrict__ in, T* __restrict__ out, int64_t nside, int order,
    int tb, int log_tb, int tiles_per_face, int64_t nrows, int rows_per_chunk,
    Side src, Side dst) {
  constexpr int kPerThread = PT;
  constexpr int kThreads = TBMAX * TBMAX / kPerThread;
  __shared__ T lds[2][TBMAX * TBMAX];

  const int ntiles = gridDim.x;
  const int tile = xcd_local_index(blockIdx.x, ntiles);
  TileGeom g;
  g.face = tile / tiles_per_face;
  const int b = tile - g.face * tiles_per_face;
  g.tb = tb;
  g.log_tb = log_tb;
  g.x0 = compress_bits(b) << log_tb;
  g.y0 = compress_bits(b >> 1) << log_tb;
  g.nest_base = ((int64_t)g.face << (2 * order)) + ((int64_t)b << (2 * log_tb));
  const int mask = (int)nside - 1;

  // Corner of the tile's image block in each flat layout.
  int sfx = 0, sfy = 0, dfx = 0, dfy = 0;
  if (src.scheme == kFlat) {
    int ax = g.x0, ay = g.y0, bx = g.x0 + tb - 1, by = g.y0 + tb - 1;
    correct(ax, ay, src.flip, src.k, mask);
    correct(bx, by, src.flip, src.k, mask);
    sfx = min(ax, bx);
    sfy = min(ay, by);
  }
  if (dst.scheme == kFlat) {
    int ax = g.x0, ay = g.y0, bx = g.x0 + tb - 1, by = g.y0 + tb - 1;
    correct(ax, ay, dst.flip, dst.k, mask);
    correct(bx, by, dst.flip, dst.k, mask);
    dfx = min(ax, bx);
    dfy = min(ay, by);
  }

  const int npt = tb * tb;
  int src_q[kPerThread], dst_q[kPerThread];
  int src_i[kPerThread], dst_i[kPerThread];  // < 2^31 for nside <= 8192
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    const int e = threadIdx.x + i * kThreads;
    src_q[i] = -1;
    dst_q[i] = -1;
    src_i[i] = 0;
    dst_i[i] = 0;
    if (e < npt) {
      int lx, ly;
      int64_t gi;
      element(src, g, e, nside, order, sfx, sfy, lx, ly, gi);
      src_q[i] = spread_bits(lx) | (spread_bits(ly) << 1);
      src_i[i] = (int)gi;
      element(dst, g, e, nside, order, dfx, dfy, lx, ly, gi);
      dst_q[i] = spread_bits(lx) | (spread_bits(ly) << 1);
      dst_i[i] = (int)gi;
    }
  }

  const int64_t npix = 12 * nside * nside;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk;
  if (r1 > nrows) r1 = nrows;
  int buf = 0;
  for (int64_t r = r0; r < r1; ++r) {
    const T* src_row = in + r * npix;
    T* dst_row = out + r * npix;
    T v[kPerThread];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
      if (src_q[i] >= 0) v[i] = NTL ? __builtin_nontemporal_load(src_row + src_i[i]) : src_row[src_i[i]];
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
      if (src_q[i] >= 0) lds[buf][src_q[i]] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPerThread; ++i)
      if (dst_q[i] >= 0) {
        if constexpr (NTS)
          __builtin_nontemporal_store(lds[buf][dst_q[i]], dst_row + dst_i[i]);
        else
          dst_row[dst_i[i]] = lds[buf][dst_q[i]];
      }
    buf ^= 1;
  }
}

// The same tiles with loading and storing split between the halves of the
// work-group: waves 0-7 load a row and store it into LDS, waves 8-15 read
// the previous row from LDS and store it to memory, one row per step, with
// one barrier per step (double-buffered LDS).  vmcnt counts a wave's loads
// and stores in issue order, so a wave that both loads and stores waits, to
// place a row's pixels in LDS, for its previous row's stores too; here the
// loading waves issue no stores and the storing waves never wait on memory.
// Whole tiles only (tb = TB): every thread moves kPer pixels per row.
template <typename T, int TB, bool NTL, bool NTS>
__global__ __launch_bounds__(1024) void remap_tiles_split(const T* __restrict__ in, T* __restrict__ out,
                                                        int64_t nside, int order, int tiles_per_face,
                                                        int64_t nrows, int rows_per_chunk, Side src, Side dst) {
  constexpr int kHalf = 512;
  constexpr int kPer = TB * TB / kHalf;
  __shared__ T lds[2][TB * TB];
  const int tile = xcd_local_index(blockIdx.x, gridDim.x);
  TileGeom g;
  g.face = tile / tiles_per_face;
  const int b = tile - g.face * tiles_per_face;
  int log_tb = 0;
  while ((1 << log_tb) < TB) ++log_tb;
  g.tb = TB;
  g.log_tb = log_tb;
  g.x0 = compress_bits(b) << log_tb;
  g.y0 = compress_bits(b >> 1) << log_tb;
  g.nest_base = ((int64_t)g.face << (2 * order)) + ((int64_t)b << (2 * log_tb));
  const int mask = (int)nside - 1;
  // wave-uniform role
  const bool loader = __builtin_amdgcn_readfirstlane(threadIdx.x) < kHalf;
  const Side side = loader ? src : dst;
  int fx0 = 0, fy0 = 0;
  if (side.scheme == kFlat) {
    int ax = g.x0, ay = g.y0, bx = g.x0 + TB - 1, by = g.y0 + TB - 1;
    correct(ax, ay, side.flip, side.k, mask);
    correct(bx, by, side.flip, side.k, mask);
    fx0 = min(ax, bx);
    fy0 = min(ay, by);
  }
  const int t = threadIdx.x & (kHalf - 1);
  // LDS slots (< 2^14) two to a register
  int gidx[kPer];
  uint32_t slot2[kPer / 2] = {};
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    int lx, ly;
    int64_t gi;
    element(side, g, t + kHalf * i, nside, order, fx0, fy0, lx, ly, gi);
    slot2[i / 2] |= (uint32_t)(spread_bits(lx) | (spread_bits(ly) << 1)) << (16 * (i & 1));
    gidx[i] = (int)gi;
  }
  auto slot = [&](int i) { return (slot2[i / 2] >> (16 * (i & 1))) & 0xffff; };
  const int64_t npix = 12 * nside * nside;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk;
  if (r1 > nrows) r1 = nrows;
  const int nsteps = (int)(r1 - r0) + 1;
  for (int s = 0; s < nsteps; ++s) {
    if (loader) {
      if (s + 1 < nsteps) {
        const T* row = in + (r0 + s) * npix;
        T v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) v[i] = NTL ? __builtin_nontemporal_load(row + gidx[i]) : row[gidx[i]];
#pragma unroll
        for (int i = 0; i < kPer; ++i) lds[s & 1][slot(i)] = v[i];
      }
    } else if (s > 0) {
      T* row = out + (r0 + s - 1) * npix;
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const T v = lds[(s - 1) & 1][slot(i)];
        if constexpr (NTS)
          __builtin_nontemporal_store(v, row + gidx[i]);
        else
          row[gidx[i]] = v;
      }
    }
    __syncthreads();
  }
}

template <typename T, int TB, bool NTL, bool NTS>
int launch_split(const void* in, void* out, int64_t nside, int64_t nrows, Side src, Side dst,
                 hipStream_t stream, int rows_per_wg) {
  int order = 0;
  while ((int64_t(1) << order) < nside) ++order;
  const int tiles_per_face = (int)((nside / TB) * (nside / TB));
  const int ntiles = 12 * tiles_per_face;
  int64_t chunks = (nrows + rows_per_wg - 1) / rows_per_wg;
  if (chunks > 65535) chunks = 65535;  // grid y limit
  const int rows_per_chunk = (int)((nrows + chunks - 1) / chunks);
  chunks = (nrows + rows_per_chunk - 1) / rows_per_chunk;
  dim3 grid(ntiles, (unsigned)chunks);
  remap_tiles_split<T, TB, NTL, NTS><<<grid, 1024, 0, stream>>>(static_cast<const T*>(in), static_cast<T*>(out),
                                                                nside, order, tiles_per_face, nrows, rows_per_chunk,
                                                                src, dst);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

// NEST -> RING through tiles that follow the rings.  A square tile's ring
// segments average TB/2 pixels, and a RING destination written in such short
// runs is slow (nest2ring 3.3 TB/s at nside 1024, against 5.2 TB/s for the
// kinds whose destination is NEST or flat).  Here a tile is a parallelogram
// of one face: the BS rings (anti-diagonals s = ix + iy) in [s0, s0 + BS),
// each cut to the BW pixels with ix in [x0, x0 + BW).  Each of its rings is
// one contiguous RING run of up to BW pixels (two where the ring wraps round
// the pixel-index origin of face 4), written with non-temporal stores: with
// runs of 1-2 KB few lines are partial, and the stores then no longer need
// the L2 to merge them (plain stores: 3.4 TB/s).  The NEST side is read as
// whole 8 x 8 Morton blocks (256 contiguous bytes at 4-byte pixels, one
// wavefront each): for every column of blocks the band meets BS / 8 + 1
// blocks, of which the first and last lie partly in the neighbouring bands
// and are masked.  Along ix a face holds nside / BW tiles; along s, the band
// of an x-range covers s in [x0, x0 + BW + nside - 2], i.e. (BW + nside) / BS
// tiles, the ones at either end only partly filled.  LDS holds the tile as
// BS rows of BW pixels with a pitch of BW + 8: an 8 x 8 block's 64 pixels
// then fall in 64 distinct banks.  Measured at [64, npix] float32:
// 512 x 32 tiles 1.47 ms at nside 1024 (256 x 64: 1.56 ms; 1024 x 16 spills
// registers: 2.3 ms); at nside 512, where 512 x 32 tiles are half empty,
// 256 x 64 tiles take 0.40 ms against 0.46 ms.
constexpr int kBandThreads = 1024;

// Raw buffer load of one 1-, 2- or 4-byte pixel (gfx9 descriptor, bounds
// checked against the descriptor's byte count): the address is a base in
// scalar registers, a wave-uniform scalar offset and one per-lane offset.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pixel_rsrc(const void* base, int nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, nbytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T pixel_load(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  if constexpr (sizeof(T) == 4)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
  else if constexpr (sizeof(T) == 2)
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0));
  else
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0));
}

// Non-temporal raw buffer store of one pixel (cache policy 2: nt).
template <typename T>
__device__ __forceinline__ void pixel_store_nt(T v, __amdgpu_buffer_rsrc_t r, int voff) {
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, 0, 2);
  else if constexpr (sizeof(T) == 2)
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, voff, 0, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b8(__builtin_bit_cast(uint8_t, v), r, voff, 0, 2);
}

__device__ __forceinline__ uint32_t band_slot(const uint32_t* q, int i) {
  return (q[i / 2] >> (16 * (i & 1))) & 0xffff;
}

// FLAT: the source is a flat layout (orientation flip, k) instead of NEST.
// Its pixel (ix, iy) lies at row y', column x' of the face's block, with
// (x', y') = correct(ix, iy); depending on the orientation a row y' is one
// row iy or one column ix of the face, and either way the band holds one run
// of at most BS consecutive pixels of it (rings s in [s0, s0 + BS)).  Each
// such run is read by BS consecutive lanes (slot e: run e / BS, pixel e % BS
// of it); rows meet the band for iy in [s0 - x0 - BW + 1, s0 + BS - 1 - x0]
// (BW + BS - 1 of them), columns for ix in [x0, x0 + BW).
template <typename T, int BW, int BS, bool FLAT = false>
__global__ __launch_bounds__(kBandThreads) void nest2ring_bands(
    const T* __restrict__ in, T* __restrict__ out, int64_t nside, int xblocks, int sblocks,
    int64_t nrows, int rows_per_chunk, int flip = 0, int rot = 0) {
  constexpr int kPitch = BW + 8;
  constexpr int kBlocksPerCol = BS / 8 + 1;
  constexpr int kRingPer = BS * BW / kBandThreads;                   // RING pixels per thread
  constexpr int kRuns = BW + BS - 1;                                 // flat runs that meet the band
  constexpr int kFlatPer = ((kRuns * BS + kBandThreads - 1) / kBandThreads + 1) / 2 * 2;
  constexpr int kNestPer = FLAT ? kFlatPer : (BW / 8) * kBlocksPerCol * 64 / kBandThreads;  // source slots per thread
  static_assert(kRingPer % 2 == 0 && kNestPer % 2 == 0 && BS * kPitch < 0xffff && 64 % BS == 0,
                "band shape");
  __shared__ T lds[2][BS * kPitch];
  const int n = (int)nside;
  const int tile = xcd_local_index(blockIdx.x, gridDim.x);
  const int per_face = xblocks * sblocks;
  const int face = tile / per_face;
  const int rem = tile - face * per_face;
  const int xb = rem / sblocks;
  const int x0 = xb * BW;
  const int s0 = x0 + (rem - xb * sblocks) * BS;

  // Source: NEST, whole 8 x 8 Morton blocks.  Block (bx, by) covers
  // s in [8 (bx + by), 8 (bx + by) + 14]; it meets the band for
  // bx + by in [s0 / 8 - 1, s0 / 8 + BS / 8 - 1].
  // LDS slots (< BS * kPitch) are kept two to a register, 0xffff marking a
  // slot outside the band: 1024 threads leave 128 VGPRs each.
  uint32_t src_q[kNestPer / 2] = {}, dst_q[kRingPer / 2] = {};
  int src_i[kNestPer];
  const int face_base = face * n * n;  // < 2^31 for nside <= 8192
  if constexpr (FLAT) {
    const int mask = n - 1;
    // does a row y' of the flat block follow iy (one row of the face)?
    int ax = 0, ay = 0, bx1 = 1, by1 = 0;
    correct(ax, ay, flip, rot, mask);
    correct(bx1, by1, flip, rot, mask);
    const bool rows_follow_iy = ay == by1;
#pragma unroll
    for (int i = 0; i < kNestPer; ++i) {
      const int e = threadIdx.x + i * kBandThreads;
      const int run = e / BS;
      const int t = e - run * BS;  // ring s0 + t of the band
      int ix, iy;
      bool ok;
      if (rows_follow_iy) {
        iy = s0 - x0 - BW + 1 + run;
        ix = s0 + t - iy;
        ok = run < kRuns && iy >= 0 && iy < n && ix >= x0 && ix < x0 + BW;
      } else {
        ix = x0 + run;
        iy = s0 + t - ix;
        ok = run < BW && iy >= 0 && iy < n;
      }
      src_q[i / 2] |= (uint32_t)(ok ? t * kPitch + (ix - x0) : 0xffff) << (16 * (i & 1));
      int fx = ix, fy = iy;
      if (ok) correct(fx, fy, flip, rot, mask);
      src_i[i] = ok ? face_base + fy * n + fx : 0;
    }
  } else
#pragma unroll
  for (int i = 0; i < kNestPer; ++i) {
    const int e = threadIdx.x + i * kBandThreads;
    const int blk = e >> 6;
    const int l = e & 63;
    const int col = blk / kBlocksPerCol;
    const int bx = (x0 >> 3) + col;
    const int by = (s0 >> 3) - 1 - bx + (blk - col * kBlocksPerCol);
    const int ix = 8 * bx + compress_bits(l);
    const int iy = 8 * by + compress_bits(l >> 1);
    const int t = ix + iy - s0;
    const bool ok = by >= 0 && by < (n >> 3) && t >= 0 && t < BS;
    src_q[i / 2] |= (uint32_t)(ok ? t * kPitch + (ix - x0) : 0xffff) << (16 * (i & 1));
    src_i[i] = ok ? face_base + (((spread_bits(bx) | (spread_bits(by) << 1)) << 6) | l) : 0;
  }
  // Destination: RING, ring t of the band, pixel p along it (increasing ix).
  int dst_i[kRingPer];
#pragma unroll
  for (int i = 0; i < kRingPer; ++i) {
    const int e = threadIdx.x + i * kBandThreads;
    const int t = e / BW;
    const int p = e - t * BW;
    const int ix = x0 + p;
    const int iy = s0 + t - ix;
    const bool ok = iy >= 0 && iy < n;
    dst_q[i / 2] |= (uint32_t)(ok ? t * kPitch + p : 0xffff) << (16 * (i & 1));
    dst_i[i] = ok ? (int)xyf2ring(ix, iy, face, nside) : 0;
  }

  const int64_t npix = 12 * nside * nside;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk;
  if (r1 > nrows) r1 = nrows;
  int buf = 0;
  for (int64_t r = r0; r < r1; ++r) {
    const T* src_row = in + r * npix;
    T* dst_row = out + r * npix;
    T v[kNestPer];
#pragma unroll
    for (int i = 0; i < kNestPer; ++i)
      if (band_slot(src_q, i) != 0xffff) v[i] = src_row[src_i[i]];
#pragma unroll
    for (int i = 0; i < kNestPer; ++i) {
      const uint32_t q = band_slot(src_q, i);
      if (q != 0xffff) lds[buf][q] = v[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kRingPer; ++i) {
      const uint32_t q = band_slot(dst_q, i);
      if (q != 0xffff) __builtin_nontemporal_store(lds[buf][q], dst_row + dst_i[i]);
    }
    buf ^= 1;
  }
}

// nest2ring_bands for 1- and 2-byte pixels.  There the kernel above
// computes a 64-bit address per pixel and spills them (its pixel loads and
// stores then wait on scratch reloads); this one addresses the NEST source
// through a per-face buffer descriptor (the block base wave-uniform in a
// scalar register, the lane's pixel as the vector offset) and the RING
// destination through 32-bit offsets from the band's lowest RING index, and
// requests row r + 1 before row r is written out.  nside 1024, [64, npix]:
// int16 2.37 -> 1.04 ms, uint8 2.02 -> 0.96 ms (bit-exact).  For 4-byte
// pixels it was slower on one box and faster on another (float32 1.48 ->
// 1.60 / 1.48 -> 1.42 ms), and slower with few rows per chunk (1024 x 7:
// 0.23 -> 0.27 ms), so 4-byte pixels keep the kernel above.
template <typename T, int BW, int BS>
__global__ __launch_bounds__(kBandThreads) void nest2ring_bands_narrow(
    const T* __restrict__ in, T* __restrict__ out, int64_t nside, int xblocks, int sblocks,
    int64_t nrows, int rows_per_chunk) {
  constexpr int kPitch = BW + 8;
  constexpr int kBlocksPerCol = BS / 8 + 1;
  constexpr int kRingPer = BS * BW / kBandThreads;                   // RING pixels per thread
  constexpr int kNestPer = (BW / 8) * kBlocksPerCol * 64 / kBandThreads;  // NEST slots per thread
  static_assert(kRingPer % 2 == 0 && kNestPer % 2 == 0 && BS * kPitch < 0xffff, "band shape");
  __shared__ T lds[2][BS * kPitch];
  const int n = (int)nside;
  const int tile = xcd_local_index(blockIdx.x, gridDim.x);
  const int per_face = xblocks * sblocks;
  const int face = tile / per_face;
  const int rem = tile - face * per_face;
  const int xb = rem / sblocks;
  const int x0 = xb * BW;
  const int s0 = x0 + (rem - xb * sblocks) * BS;

  // Source: NEST, whole 8 x 8 Morton blocks.  Block (bx, by) covers
  // s in [8 (bx + by), 8 (bx + by) + 14]; it meets the band for
  // bx + by in [s0 / 8 - 1, s0 / 8 + BS / 8 - 1].  Slot e of the source
  // side is pixel l = e % 64 of block e / 64, so a wavefront reads one
  // whole block: the block's base index is wave-uniform (kept in scalar
  // registers) and a lane adds its own l.  Lanes whose pixel lies outside
  // the band still load it (every block inside the face is in the map; a
  // block outside the face reads the row's first 64 pixels), and skip the
  // LDS store: loads under a branch would make the compiler wait for every
  // load and store in flight, the previous row's stores included.
  // LDS slots (< BS * kPitch) are kept two to a register, 0xffff marking a
  // slot outside the band: 1024 threads leave 128 VGPRs each.
  uint32_t src_q[kNestPer / 2] = {}, dst_q[kRingPer / 2] = {};
  int src_off[kNestPer];  // byte offset of the block in its face (< 2^31)
  const int face_base = face * n * n;  // < 2^31 for nside <= 8192
  const int l = threadIdx.x & 63;
  const int lx = compress_bits(l), ly = compress_bits(l >> 1);
#pragma unroll
  for (int i = 0; i < kNestPer; ++i) {
    const int blk = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6) + i * (kBandThreads / 64));
    const int col = blk / kBlocksPerCol;
    const int bx = (x0 >> 3) + col;
    const int by = (s0 >> 3) - 1 - bx + (blk - col * kBlocksPerCol);
    const bool in_face = by >= 0 && by < (n >> 3);
    const int ix = 8 * bx + lx;
    const int t = ix + 8 * by + ly - s0;
    const bool ok = in_face && t >= 0 && t < BS;
    src_q[i / 2] |= (uint32_t)(ok ? t * kPitch + (ix - x0) : 0xffff) << (16 * (i & 1));
    src_off[i] = in_face ? ((spread_bits(bx) | (spread_bits(by) << 1)) << 6) * (int)sizeof(T) : 0;
  }
  // Destination: RING, ring t of the band, pixel p along it (increasing ix).
  // The band's pixels lie in at most BS consecutive rings, so they are
  // addressed as byte offsets from the band's lowest RING index (band_lo,
  // found once per work-group): 32-bit buffer offsets instead of a 64-bit
  // address per pixel.
  __shared__ int band_lo;
  if (threadIdx.x == 0) band_lo = 0x7fffffff;
  int dst_i[kRingPer];
  int lo = 0x7fffffff;
#pragma unroll
  for (int i = 0; i < kRingPer; ++i) {
    const int e = threadIdx.x + i * kBandThreads;
    const int t = e / BW;
    const int p = e - t * BW;
    const int ix = x0 + p;
    const int iy = s0 + t - ix;
    const bool ok = iy >= 0 && iy < n;
    dst_q[i / 2] |= (uint32_t)(ok ? t * kPitch + p : 0xffff) << (16 * (i & 1));
    dst_i[i] = ok ? (int)xyf2ring(ix, iy, face, nside) : 0;
    if (ok) lo = min(lo, dst_i[i]);
  }
  __syncthreads();
  atomicMin(&band_lo, lo);
  __syncthreads();
  lo = __builtin_amdgcn_readfirstlane(band_lo);
#pragma unroll
  for (int i = 0; i < kRingPer; ++i)
    dst_i[i] = band_slot(dst_q, i) != 0xffff ? (dst_i[i] - lo) * (int)sizeof(T) : 0;

  const int64_t npix = 12 * nside * nside;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per_chunk;
  int64_t r1 = r0 + rows_per_chunk;
  if (r1 > nrows) r1 = nrows;
  // two register sets: row r + 1 is requested before row r is written out,
  // so its loads are in flight while row r passes through LDS and leaves
  // (the reload past the last row reads the last row again, unused)
  T va[kNestPer], vb[kNestPer];
  const int voff = l * (int)sizeof(T);
  const int face_bytes = n * n * (int)sizeof(T);  // < 2^31 for nside <= 8192
  auto load_row = [&](T (&v)[kNestPer], int64_t r) {
    const __amdgpu_buffer_rsrc_t face_rsrc = pixel_rsrc(in + (r < r1 ? r : r1 - 1) * npix + face_base, face_bytes);
#pragma unroll
    for (int i = 0; i < kNestPer; ++i) v[i] = pixel_load<T>(face_rsrc, voff, src_off[i]);
  };
  auto put_row = [&](const T (&v)[kNestPer], int b) {
#pragma unroll
    for (int i = 0; i < kNestPer; ++i) {
      const uint32_t q = band_slot(src_q, i);
      if (q != 0xffff) lds[b][q] = v[i];
    }
  };
  // the band's bytes run from RING index lo to at most BS + 1 rings of
  // 4 nside pixels on (a few MB); the descriptor allows all of that
  const int band_bytes = (int)min<int64_t>((npix - lo) * (int64_t)sizeof(T), 0x7fffffff);
  auto emit_row = [&](int b, int64_t r) {
    const __amdgpu_buffer_rsrc_t band_rsrc = pixel_rsrc(out + r * npix + lo, band_bytes);
#pragma unroll
    for (int i = 0; i < kRingPer; ++i) {
      const uint32_t q = band_slot(dst_q, i);
      if (q != 0xffff) pixel_store_nt<T>(lds[b][q], band_rsrc, dst_i[i]);
    }
  };
  if (r0 >= r1) return;
  load_row(va, r0);
  int64_t r = r0;
  for (; r + 1 < r1; r += 2) {
    load_row(vb, r + 1);
    put_row(va, 0);
    __syncthreads();
    emit_row(0, r);
    load_row(va, r + 2);
    put_row(vb, 1);
    __syncthreads();
    emit_row(1, r + 1);
  }
  if (r < r1) {
    put_row(va, 0);
    __syncthreads();
    emit_row(0, r);
  }
}

template <typename T, int BW, int BS, bool FLAT = false>
int launch_bands(const void* in, void* out, int64_t nside, int64_t nrows, hipStream_t stream, int flip = 0,
                 int rot = 0) {
  if (nside < BW) HPX_FAIL(HPX_ERR_UNSUPPORTED, "band tiles need nside >= %d", BW);
  const int xblocks = (int)(nside / BW);
  const int sblocks = (int)((BW + nside) / BS);
  const int ntiles = 12 * xblocks * sblocks;
  // Enough work-groups to fill 256 CUs (one ~132 KB work-group each) about
  // six times: two row chunks at nside 1024 (one chunk: 1.49 ms, three:
  // 1.48 ms, two: 1.47 ms).
  const int64_t want = 1536;
  int64_t chunks = (want + ntiles - 1) / ntiles;
  if (chunks > nrows) chunks = nrows;
  if (chunks < 1) chunks = 1;
  const int rows_per_chunk = (int)((nrows + chunks - 1) / chunks);
  chunks = (nrows + rows_per_chunk - 1) / rows_per_chunk;
  dim3 grid(ntiles, (unsigned)chunks);
  static_assert(!(FLAT && sizeof(T) < 4), "flat-source bands are built for 4-byte pixels");
  if constexpr (sizeof(T) < 4)
    nest2ring_bands_narrow<T, BW, BS><<<grid, kBandThreads, 0, stream>>>(
        static_cast<const T*>(in), static_cast<T*>(out), nside, xblocks, sblocks, nrows, rows_per_chunk);
  else
    nest2ring_bands<T, BW, BS, FLAT><<<grid, kBandThreads, 0, stream>>>(
        static_cast<const T*>(in), static_cast<T*>(out), nside, xblocks, sblocks, nrows,
        rows_per_chunk, flip, rot);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

// One 16-byte vector per thread, no loop: a short-lived work-group per 4 KB,
// non-temporal loads and stores.  Measured on MI355X (4 GiB): 6.48 TB/s;
// plain loads and stores 6.20, non-temporal stores only 6.33, 512- or
// 1024-thread work-groups 5.9-6.2, two or four vectors per thread 5.6-6.0,
// a grid-stride loop over 16384 work-groups 4.9 TB/s.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy16(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                              int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}

__global__ void copy1(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// flat -> flat for the orientation pairs whose composite keeps the axes
// (identity, x mirrored, y mirrored, half turn): pixel (x, y) of a face's
// block goes to (rx ? n-1-x : x, ry ? n-1-y : y) of the same face, so every
// 16 bytes of a row land as 16 bytes of one destination row, element order
// reversed when rx.  One vector per thread, as the copy kernel (copy16):
// non-temporal loads and stores, short-lived 256-thread work-groups; grid.x
// covers one map row's vectors, grid.y the batch rows of this launch.
template <typename T>
__global__ __launch_bounds__(256) void flat_rows16(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                   int log_n, int rx, int ry) {
  constexpr int kPer = 16 / (int)sizeof(T);
  constexpr int kLogPer = kPer == 16 ? 4 : kPer == 8 ? 3 : kPer == 4 ? 2 : 1;
  const int64_t vec_per_row = (int64_t)3 << (2 * log_n + 2 - kLogPer);  // 12 n^2 / kPer
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= vec_per_row) return;
  const int n = 1 << log_n;
  const int64_t p = v << kLogPer;  // first pixel of the vector
  const int64_t face = p >> (2 * log_n);
  const int q = (int)(p & (((int64_t)1 << (2 * log_n)) - 1));
  const int y = q >> log_n, x = q & (n - 1);
  const int yd = ry ? n - 1 - y : y;
  const int xd = rx ? n - kPer - x : x;
  const int64_t vd = ((face << (2 * log_n)) + ((int64_t)yd << log_n) + xd) >> kLogPer;
  const int64_t row = (int64_t)blockIdx.y * vec_per_row;
  u32x4 val = __builtin_nontemporal_load(in + row + v);
  if (rx) {
    T e[kPer];
    __builtin_memcpy(e, &val, 16);
#pragma unroll
    for (int i = 0; i < kPer / 2; ++i) {
      const T t = e[i];
      e[i] = e[kPer - 1 - i];
      e[kPer - 1 - i] = t;
    }
    __builtin_memcpy(&val, e, 16);
  }
  __builtin_nontemporal_store(val, out + row + vd);
}

// The axis-keeping composite of two flat orientations, if it is one:
// returns true with the mirror flags of (x, y) -> (x', y').
bool flat_pair_keeps_axes(const Side& src, const Side& dst, int64_t nside, int& rx, int& ry) {
  const int mask = (int)nside - 1;
  auto comp = [&](int x, int y, int& xo, int& yo) {
    uncorrect(x, y, src.flip, src.k, mask);
    correct(x, y, dst.flip, dst.k, mask);
    xo = x;
    yo = y;
  };
  int x00, y00, x10, y10, x01, y01;
  comp(0, 0, x00, y00);
  comp(1, 0, x10, y10);
  comp(0, 1, x01, y01);
  if (y10 != y00 || x01 != x00) return false;  // the axes are exchanged
  rx = x10 < x00;
  ry = y01 < y00;
  // every axis-keeping element of the square's symmetry group is of this
  // form; check two more points all the same
  for (int t = 0; t < 2; ++t) {
    const int x = t ? mask : 2 & mask, y = t ? 3 & mask : mask;
    int xo, yo;
    comp(x, y, xo, yo);
    if (xo != (rx ? mask - x : x) || yo != (ry ? mask - y : y)) return false;
  }
  return true;
}

// flip / k of a flat layout, as the reference wrappers derive them for the
// NEST-side (origin S, counter-clockwise) partner: flip = cw,
// k = (cw ? -o : o) & 3 (hpx_remapping_cuda.cu:415-418 with src (0, ccw)).
Side flat_side(int origin, int cw) {
  Side s;
  s.scheme = kFlat;
  s.flip = cw ? 1 : 0;
  s.k = (cw ? -origin : origin) & 3;
  return s;
}

// want: work-groups to aim for (tiles x row chunks); more chunks mean fewer
// rows per work-group and the tile's index arithmetic done more often.
// rows_per_wg > 0 instead fixes the rows per work-group; min_rows keeps at
// least that many rows per work-group (fewer chunks) where want asks for more.
template <typename T, int TBMAX, int PT, bool NTL, bool NTS>
int launch_tiles(const void* in, void* out, int64_t nside, int64_t nrows, Side src, Side dst,
                 hipStream_t stream, int64_t want, int rows_per_wg = 0, int min_rows = 1) {
  constexpr int kThreads = TBMAX * TBMAX / PT;
  int order = 0;
  while ((int64_t(1) << order) < nside) ++order;
  const int tb = nside < TBMAX ? (int)nside : TBMAX;
  int log_tb = 0;
  while ((1 << log_tb) < tb) ++log_tb;
  const int tiles_per_face = (int)((nside / tb) * (nside / tb));
  const int ntiles = 12 * tiles_per_face;
  int64_t chunks = rows_per_wg > 0 ? (nrows + rows_per_wg - 1) / rows_per_wg : (want + ntiles - 1) / ntiles;
  if (min_rows > 1 && chunks > nrows / min_rows) chunks = nrows / min_rows;
  if (chunks > nrows) chunks = nrows;
  if (chunks > 65535) chunks = 65535;  // grid y limit
  if (chunks < 1) chunks = 1;
  const int rows_per_chunk = (int)((nrows + chunks - 1) / chunks);
  chunks = (nrows + rows_per_chunk - 1) / rows_per_chunk;
  dim3 grid(ntiles, (unsigned)chunks);
  remap_tiles<T, TBMAX, PT, NTL, NTS><<<grid, kThreads, 0, stream>>>(
      static_cast<const T*>(in), static_cast<T*>(out), nside, order, tb, log_tb,
      tiles_per_face, nrows, rows_per_chunk, src, dst);
  HPX_LAUNCH_CHECK();
  return HPX_OK;
}

template <typename T>
int launch(const void* in, void* out, int64_t nside, int64_t nrows, Side src, Side dst,
           hipStream_t stream) {
  const bool ring = src.scheme == kRing || dst.scheme == kRing;
  const bool nest_flat = (src.scheme == kNest && dst.scheme == kFlat) || (src.scheme == kFlat && dst.scheme == kNest);
  // flat -> flat keeping the axes: whole rows, at the copy kernel's rate
  // (nside 1024, [64, npix] float32, origin S ccw -> W cw: 1.19 ms with the
  // split 64 x 64 tiles below)
  int rx = 0, ry = 0;
  if (src.scheme == kFlat && dst.scheme == kFlat && nside * (int64_t)sizeof(T) >= 16 &&
      ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0 &&
      flat_pair_keeps_axes(src, dst, nside, rx, ry)) {
    int log_n = 0;
    while ((int64_t(1) << log_n) < nside) ++log_n;
    const int64_t vec_per_row = 12 * nside * nside * (int64_t)sizeof(T) / 16;
    for (int64_t r0 = 0; r0 < nrows; r0 += 65535) {
      const int64_t nr = std::min<int64_t>(nrows - r0, 65535);
      const dim3 grid((unsigned)((vec_per_row + 255) / 256), (unsigned)nr);
      flat_rows16<T><<<grid, 256, 0, stream>>>(static_cast<const u32x4*>(in) + r0 * vec_per_row,
                                               static_cast<u32x4*>(out) + r0 * vec_per_row, log_n, rx, ry);
      HPX_LAUNCH_CHECK();
    }
    return HPX_OK;
  }
  if constexpr (sizeof(T) <= 4) {
    if (src.scheme == kNest && dst.scheme == kRing) {
      if (nside >= 1024) return launch_bands<T, 512, 32>(in, out, nside, nrows, stream);
      if (nside == 512) return launch_bands<T, 256, 64>(in, out, nside, nrows, stream);
    }
  }
  if constexpr (sizeof(T) == 4) {
    // flat -> RING with 4-byte pixels from nside 512: the ring-following
    // bands of NEST -> RING with a flat source (flat2ring at nside 1024,
    // [64, npix] float32: 1.45-1.52 ms for the four origins, against 1.95 ms
    // with 128 x 128 tiles whose RING runs average 64 pixels, and 2.20 ms
    // with the split 64 x 64 tiles below; nside 2048: 0.95 -> 0.73 ms)
    if (src.scheme == kFlat && dst.scheme == kRing) {
      if (nside >= 1024) return launch_bands<T, 512, 32, true>(in, out, nside, nrows, stream, src.flip, src.k);
      if (nside == 512) return launch_bands<T, 256, 64, true>(in, out, nside, nrows, stream, src.flip, src.k);
    }
  }
  // Every other kind from nside 64 (RING kinds from 128): 64 x 64 tiles whose
  // loads and stores are done by separate wave halves (remap_tiles_split),
  // 4 rows per work-group for the NEST / flat kinds and 8 for the RING kinds,
  // non-temporal loads and stores between NEST and flat and non-temporal
  // stores of a NEST / flat destination written from RING.  Measured at
  // nside 1024, [64, npix], against 64 x 64 tiles with both roles in every
  // wave (NEST / flat kinds) or 128 x 128 tiles (RING kinds):
  //   float32 nest2flat 1.20 -> 1.11 ms, flat2flat 1.26 -> 1.18, ring2nest
  //   1.25 -> 1.25, ring2flat 1.32 -> 1.29 (nside 256: ring2nest 0.090 ->
  //   0.074, ring2flat 0.114 -> 0.074 ms);
  //   int16 nest2flat 1.47 -> 0.82, ring2nest 1.87 -> 0.84, flat2ring
  //   2.04 -> 1.01; uint8 1.40 -> 0.73, 1.54 -> 0.69; int64 4.28 -> 2.28,
  //   ring2nest 4.71 -> 2.40, nest2ring 5.04 -> 3.69 ms.
  // (The earlier shapes: 4 rows per work-group beat 64 rows per work-group
  // for the NEST / flat kinds, nside 512: 2.9 TB/s with one row, 5.6 with 4.)
  if (ring && nside >= 128) {
    if (src.scheme == kRing) return launch_split<T, 64, false, true>(in, out, nside, nrows, src, dst, stream, 8);
    return launch_split<T, 64, false, false>(in, out, nside, nrows, src, dst, stream, 8);
  }
  if (!ring && nside >= 64) {
    if (nest_flat) return launch_split<T, 64, true, true>(in, out, nside, nrows, src, dst, stream, 4);
    return launch_split<T, 64, false, false>(in, out, nside, nrows, src, dst, stream, 4);
  }
  // small maps: tiles of the whole face side with both roles in every wave
  if (!ring && sizeof(T) <= 4) {
    if (nest_flat) return launch_tiles<T, 64, 4, true, true>(in, out, nside, nrows, src, dst, stream, 0, 4);
    return launch_tiles<T, 64, 4, false, false>(in, out, nside, nrows, src, dst, stream, 0, 4);
  }
  if (nest_flat) return launch_tiles<T, 64, 16, true, true>(in, out, nside, nrows, src, dst, stream, 2048);
  return launch_tiles<T, 64, 16, false, false>(in, out, nside, nrows, src, dst, stream, 2048);
}

}  // namespace

extern "C" int hpx_remap(int kind, int64_t nside, int origin, int clockwise, int dst_origin,
                         int dst_clockwise, const void* in, void* out, int64_t nrows,
                         int elem_bytes, hipStream_t stream) {
  if (nside < 1 || (nside & (nside - 1)) != 0)
    HPX_FAIL(HPX_ERR_ARG, "nside must be a power of two (got %lld)", (long long)nside);
  if (nside > 8192) HPX_FAIL(HPX_ERR_UNSUPPORTED, "nside > 8192 is not supported");
  if (nrows < 0) HPX_FAIL(HPX_ERR_ARG, "nrows must be >= 0");
  if (origin < 0 || origin > 3 || dst_origin < 0 || dst_origin > 3)
    HPX_FAIL(HPX_ERR_ARG, "The origin must be one from S, E, N, W. Stop");
  if (nrows == 0) return HPX_OK;
  if (in == nullptr || out == nullptr) HPX_FAIL(HPX_ERR_ARG, "null data pointer");
  if (in == out) HPX_FAIL(HPX_ERR_ARG, "in-place remap is not supported");

  const Side nest{kNest, 0, 0};
  const Side ring{kRing, 0, 0};
  Side src, dst;
  switch (kind) {
    case HPX_RING2NEST: src = ring; dst = nest; break;
    case HPX_NEST2RING: src = nest; dst = ring; break;
    case HPX_NEST2FLAT: src = nest; dst = flat_side(origin, clockwise); break;
    case HPX_FLAT2NEST: src = flat_side(origin, clockwise); dst = nest; break;
    case HPX_RING2FLAT: src = ring; dst = flat_side(origin, clockwise); break;
    case HPX_FLAT2RING: src = flat_side(origin, clockwise); dst = ring; break;
    case HPX_FLAT2FLAT:
      src = flat_side(origin, clockwise);
      dst = flat_side(dst_origin, dst_clockwise);
      break;
    default: HPX_FAIL(HPX_ERR_ARG, "unknown remap kind %d", kind);
  }
  switch (elem_bytes) {
    case 1: return launch<uint8_t>(in, out, nside, nrows, src, dst, stream);
    case 2: return launch<uint16_t>(in, out, nside, nrows, src, dst, stream);
    case 4: return launch<uint32_t>(in, out, nside, nrows, src, dst, stream);
    case 8: return launch<uint64_t>(in, out, nside, nrows, src, dst, stream);
    default: HPX_FAIL(HPX_ERR_ARG, "elem_bytes must be 1, 2, 4 or 8 (got %d)", elem_bytes);
  }
}

extern "C" int hpx_copy(const void* in, void* out, int64_t nbytes, hipStream_t stream) {
  if (nbytes < 0) HPX_FAIL(HPX_ERR_ARG, "nbytes must be >= 0");
  if (nbytes == 0) return HPX_OK;
  const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  const int64_t nvec = aligned ? nbytes / 16 : 0;
  for (int64_t done = 0; done < nvec;) {
    // at most 2^31 - 1 work-groups per launch (32 TiB)
    const int64_t part = std::min<int64_t>(nvec - done, int64_t(0x7fffffff) * 256);
    copy16<<<(unsigned)((part + 255) / 256), 256, 0, stream>>>(static_cast<const u32x4*>(in) + done,
                                                                static_cast<u32x4*>(out) + done, part);
    HPX_LAUNCH_CHECK();
    done += part;
  }
  const int64_t tail = nbytes - nvec * 16;
  if (tail > 0) {
    copy1<<<(unsigned)((tail + 255) / 256), 256, 0, stream>>>(
        static_cast<const uint8_t*>(in) + nvec * 16, static_cast<uint8_t*>(out) + nvec * 16, tail);
    HPX_LAUNCH_CHECK();
  }
  return HPX_OK;
}
