// libssq engine: Z2 boundary-matrix reduction on MI355X (gfx950, wave64).
//
// Pipeline for one matrix (or a block-diagonal batch of matrices), all device-resident:
//   model bit 1  edge collapse: edge list gathered on device, kept set from the host
//                collapser, coface propagation on device            (specseq.h:410-438,
//                                                                     kernel.h:190-198)
//   model bit 0  high-dimension guided clearing of max-dim cells      (kernel.h:121-292)
//   dualize      the anti-transpose by a stable radix sort (cohomology);
//                pairs map back as (b', d') -> (n-1-d', n-1-b')       (specseq.h:303-363,
//                                                                     kernel.h:521-580)
//   init passes  low / leftmost-one, apparent pairs, clearing, compression
//                                                                     (kernel.h:444-520)
//   page loop    per dimension from high to low, spectral-sequence pages of columns x
//                rows; each page runs claim/add rounds until every active column is a
//                pivot, zero or below the page's row floor            (specseq.h:592-704)
//                A round: claims, the claimants' pivot metadata, additions by one
//                wavefront per column (LDS merge), then additions of the columns that
//                outgrew a wavefront's LDS by one workgroup each (LDS bit set).
//   twist        rows of the dimension's pivots are zeroed before the next dimension
//                                                                     (specseq.h:1039-1046)
//   pairs        (row, pivot column) in row order                     (specseq.h:1182-1193)
//
// Exactness of the concurrent rounds.  A claim round does atomicMin(piv[low[c]], c) for the
// columns whose low is new; after it, every column whose low is owned by a smaller column
// is in the add list.  During an add round nobody writes piv, and a column that is piv[l]
// is never in the add list, so pivot columns are read-only while others add them.  Every
// addition adds a column to the left (p < c), which keeps all lower-left ranks and hence the
// pairs.  A chasing column keeps adding piv[low] while that is a smaller column.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <atomic>
#include <memory>
#include <thread>
#include <vector>

#include "ssq_internal.h"
#include "ssq_kernels.h"

namespace ssq {

// ------------------------------------------------------------------------------------
// device state
// ------------------------------------------------------------------------------------
constexpr int kAddWavesDefault = 4;    // wavefronts per short-addition workgroup (2, 4 or 8)
#ifndef SSQ_KTILE
#define SSQ_KTILE 1024
#endif
constexpr int kTile = SSQ_KTILE;       // LDS entries per staging buffer
constexpr int kTileBuf = kTile + kWave;
constexpr int kMaxDims = 8;            // cell dimensions 0..7

// Where the pivot column of a row is stored, written once the pivot is established so that
// an addition can fetch the pivot's id and its storage in one round trip.
struct __align__(32) PivMeta {
  int64_t off;     // start of the pivot column in the column store
  int32_t len;     // its length
  int32_t first;   // its smallest row
  int32_t second;  // its second-largest row (the largest is the row itself), -1 if none
  int32_t pad[3];
};
__device__ __forceinline__ PivMeta pivmeta_of(int64_t off, int32_t len, const int32_t* col) {
  PivMeta m;
  m.off = off;
  m.len = len;
  m.first = col[0];
  m.second = len >= 2 ? col[len - 2] : -1;
  m.pad[0] = m.pad[1] = m.pad[2] = 0;
  return m;
}

struct Dev {
  int64_t n;           // cells (rows)
  int64_t ncols;       // cells [0, ncols) have columns; the rest are rows only (<= n)
  PivMeta* pivmeta;    // row -> storage of piv[row]
  const int32_t* dims;
  const int64_t* colptr;
  int32_t* pool;       // column store: [input rows | arena]
  int64_t* off;
  int32_t* len;
  int32_t* cap;
  int32_t* low;
  uint8_t* state;
  uint8_t* removed;
  uint8_t* flag;       // scratch per-cell flag (clear / negative)
  uint32_t* piv;       // row -> column
  uint32_t* leftone;   // row -> leftmost column containing it
  int32_t* work[6];    // W0, W1, A0, A1, long list, huge list
  uint32_t* gbits;     // huge-column bit sets: one region of gspan bits per k_add_huge workgroup
  int32_t gspan;       // rows covered by one region
  uint32_t* ctr;       // [0..1] cW, [2..3] cA, [4] apparent, [5] hd-cleared
  unsigned long long* arena_used;  // ints used past arena_base
  int64_t arena_base, arena_cap;
  uint32_t* lvl_cnt;   // slots taken per level 1..3
  int32_t lvl_cap[4];
  uint32_t lvl_slots[4];
  unsigned long long* stats;  // counters at cw::kStats
  uint32_t* err;
  int32_t short_max;    // largest |a|+|b| merged in a wave's LDS (<= kTile)
  int32_t long_window;  // rows covered by the long-column bit set (<= kLongBits)
};

// Word offsets in the per-call counter block (64 x uint32 in device memory).
namespace cw {
constexpr int kW = 0;          // [0..1] worklist counts (claim input), double-buffered by round
constexpr int kA = 2;          // [2..3] add-list counts
constexpr int kApparent = 4;   // apparent pairs found by the init passes
constexpr int kHdCleared = 5;  // cells cleared by guided clearing
constexpr int kLong = 6;       // long-column list count (this round)
constexpr int kHuge = 7;       // huge-column list count (this round)
constexpr int kFlags = 8;      // [8..9] input validation flags
constexpr int kHugeNext = 10;  // next huge-list column to take (this round)
constexpr int kRowStat = 16;   // [16..23] max low+1 per dim, [24..31] min first row per dim
constexpr int kStats = 32;     // 8 x u64: additions, bytes, tiled additions, dense additions, dense bytes
constexpr int kArena = 48;     // u64: arena entries used
constexpr int kLvl = 52;       // [52..55] level slot counters
constexpr int kErr = 56;       // store exhausted
constexpr int kMaxDim = 58;    // [58] max dimension, [59] bad dimension flag
constexpr int kChain = 60;     // [60] longest chain, [61] largest |a|+|b|
constexpr int kNegDims = 62;   // bit d set: some negative cell has dimension d
constexpr int kWords = 64;
}  // namespace cw

// Element loops over cells.  Workgroups are dispatched round-robin over the 8 XCDs, so with a
// grid that is a multiple of 8, workgroup b runs on XCD b % 8.  Each XCD is given one
// contiguous eighth of [0, n): the cells in flight on an XCD at a time are neighbours, and
// their scattered accesses (rows of a column, cofaces, pivot tables) stay within the few
// complexes of a block-diagonal batch that that XCD's L2 holds.  Other grids stride plainly.
constexpr int kXcds = 8;
__device__ __forceinline__ bool xcd_grid() { return gridDim.x >= kXcds && (gridDim.x % kXcds) == 0; }
__device__ __forceinline__ int64_t xcd_span(int64_t n) { return (n + kXcds - 1) / kXcds; }
__device__ __forceinline__ int64_t xs_begin(int64_t n) {
  if (!xcd_grid()) return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  return (int64_t)(blockIdx.x % kXcds) * xcd_span(n) + (int64_t)(blockIdx.x / kXcds) * blockDim.x + threadIdx.x;
}
__device__ __forceinline__ int64_t xs_step() {
  return xcd_grid() ? (int64_t)(gridDim.x / kXcds) * blockDim.x : (int64_t)gridDim.x * blockDim.x;
}
__device__ __forceinline__ int64_t xs_end(int64_t n) {
  if (!xcd_grid()) return n;
  const int64_t e = (int64_t)(blockIdx.x % kXcds + 1) * xcd_span(n);
  return e < n ? e : n;
}
#define GRID_STRIDE(i, n) \
  for (int64_t i = xs_begin(n), i##_step = xs_step(), i##_end = xs_end(n); i < i##_end; i += i##_step)

// ------------------------------------------------------------------------------------
// allocation in the column store (level slots, then the exact-size overflow arena)
// ------------------------------------------------------------------------------------
__device__ bool store_alloc(const Dev& d, int32_t need, int64_t* off, int32_t* capo) {
  for (int k = 1; k <= 3; k++) {
    if (need <= d.lvl_cap[k]) {
      const uint32_t s = atomicAdd(&d.lvl_cnt[k], 1u);
      if (s < d.lvl_slots[k]) {
        const unsigned long long o = atomicAdd(d.arena_used, (unsigned long long)d.lvl_cap[k]);
        if ((int64_t)o + d.lvl_cap[k] <= d.arena_cap) {
          *off = d.arena_base + (int64_t)o;
          *capo = d.lvl_cap[k];
          return true;
        }
        atomicOr(d.err, 1u);
        return false;
      }
    }
  }
  const int64_t sz = ((int64_t)need + 63) & ~(int64_t)63;
  const unsigned long long o = atomicAdd(d.arena_used, (unsigned long long)sz);
  if ((int64_t)o + sz <= d.arena_cap) {
    *off = d.arena_base + (int64_t)o;
    *capo = (int32_t)sz;
    return true;
  }
  atomicOr(d.err, 1u);
  return false;
}

// wave-uniform allocation by lane 0
__device__ bool wave_alloc(const Dev& d, int32_t need, int64_t* off, int32_t* capo) {
  int64_t o = 0;
  int32_t c = 0;
  int ok = 0;
  if (lane_id() == 0) ok = store_alloc(d, need, &o, &c) ? 1 : 0;
  ok = wave_bcast(ok, 0);
  o = wave_bcast64(o, 0);
  c = wave_bcast(c, 0);
  *off = o;
  *capo = c;
  return ok != 0;
}

// ------------------------------------------------------------------------------------
// setup and init passes
// ------------------------------------------------------------------------------------
// Column state over [0, ncols); with init_rows (ncols == n) also the per-row tables, which a
// prebuilt matrix brings initialised instead.
// Byte fill and byte copy as kernels.  The engine fills and copies device memory only through
// these, never through the runtime's memset and device-to-device copy calls: a caller may
// capture the engine's work in a graph, where memset nodes have given wrong results on replay.
// [0, head) and the tail go byte by byte; 16-byte words in between (head aligns dst to 16 bytes;
// a copy uses words only where src has the same alignment, else words16 = 0).
__global__ void k_fill_bytes(uint8_t* p, uint32_t v4, int64_t head, int64_t words16, int64_t bytes) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  if (t < head) p[t] = (uint8_t)v4;
  const uint4 w = make_uint4(v4, v4, v4, v4);
  uint4* q = (uint4*)(p + head);
  for (int64_t i = t; i < words16; i += step) q[i] = w;
  for (int64_t i = head + 16 * words16 + t; i < bytes; i += step) p[i] = (uint8_t)v4;
}
__global__ void k_copy_bytes(uint8_t* dst, const uint8_t* src, int64_t head, int64_t words16, int64_t bytes) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t)gridDim.x * blockDim.x;
  if (t < head) dst[t] = src[t];
  uint4* q = (uint4*)(dst + head);
  const uint4* r = (const uint4*)(src + head);
  for (int64_t i = t; i < words16; i += step) q[i] = r[i];
  for (int64_t i = head + 16 * words16 + t; i < bytes; i += step) dst[i] = src[i];
}

__global__ void k_setup(Dev d, int init_rows) {
  GRID_STRIDE(c, d.ncols) {
    const int64_t o = d.colptr[c];
    const int32_t l = (int32_t)(d.colptr[c + 1] - o);
    const bool rem = d.removed[c] != 0;
    d.off[c] = o;
    d.cap[c] = l;
    d.len[c] = rem ? 0 : l;
    d.low[c] = (rem || l == 0) ? -1 : d.pool[o + l - 1];
    d.state[c] = (rem || l == 0) ? ST_ZERO : ST_PENDING;
    if (init_rows) {
      d.piv[c] = kEmpty;
      d.leftone[c] = kEmpty;
      d.flag[c] = 0;
      d.flag[d.n + c] = 0;
    }
  }
}

// Low[col] and leftmost-one per row (kernel.h:444-457)
__global__ void k_leftone(Dev d) {
  GRID_STRIDE(c, d.n) {
    const int32_t l = d.len[c];
    const int32_t* col = d.pool + d.off[c];
    for (int32_t k = 0; k < l; k++) atomicMin(&d.leftone[col[k]], (uint32_t)c);
  }
}

// Apparent pairs and the positive-facet flag (kernel.h:459-481): the youngest facet of a
// column is positive (its column reduces to zero; flag[row] = 1); a column that is leftmost
// in its low row is an apparent pair; a column leftmost in any row is negative
// (flag[n + col] = 1).  The two flags live in separate bytes: plain stores of 1 only.
// Also collects the dimensions that hold a negative cell (one bit each, for compression).
__global__ void k_lookup(Dev d, const uint8_t* hdclear, int neg_known) {
  uint32_t apparent = 0;  // summed per wavefront: one atomic per wave, not per pair
  uint32_t negdims = 0;
  GRID_STRIDE(c, d.ncols) {
    const int32_t l = d.len[c];
    if (l == 0) continue;
    const int32_t lw = d.low[c];
    d.flag[lw] = 1;
    if (d.leftone[lw] == (uint32_t)c) {
      d.piv[lw] = (uint32_t)c;
      d.state[c] = ST_PIVOT;
      apparent++;
    }
    bool neg;
    if (neg_known) {
      neg = d.flag[d.n + c] != 0;
    } else {
      neg = false;
      const int32_t* col = d.pool + d.off[c];
      for (int32_t k = 0; k < l; k++)
        if (d.leftone[col[k]] == (uint32_t)c) {
          d.flag[d.n + c] = 1;
          neg = true;
          break;
        }
    }
    if (neg) {
      const int32_t dm = d.dims[c];
      if (dm >= 0 && dm < 32) negdims |= 1u << dm;
    }
  }
  for (int k = kWave / 2; k > 0; k >>= 1) {
    apparent += __shfl_xor(apparent, k, kWave);
    negdims |= __shfl_xor(negdims, k, kWave);
  }
  if (lane_id() == 0 && apparent) atomicAdd(&d.ctr[cw::kApparent], apparent);
  if (lane_id() == 0 && negdims) atomicOr(&d.ctr[cw::kNegDims], negdims);
  if (hdclear)
    GRID_STRIDE(c, d.ncols) if (hdclear[c]) d.flag[c] = 1;
}


// Dual init from the primal (no pass over the dual's nonzeros): for a primal cell c with
// boundary, the dual row n-1-c has its leftmost one in dual column n-1-low(c), and the dual
// column n-1-low(c) is leftmost in that row, i.e. negative (Lemma L in the dual).
__global__ void k_dual_init_from_primal(int64_t n, const int64_t* colptr, const int32_t* rows,
                                        const uint8_t* removed_primal, uint32_t* leftone, uint8_t* negflag) {
  GRID_STRIDE(c, n) {
    const int64_t e = colptr[c + 1];
    if (removed_primal[c] || e == colptr[c]) continue;
    const int32_t low = rows[e - 1];
    leftone[n - 1 - c] = (uint32_t)(n - 1 - low);
    negflag[n - 1 - low] = 1;
  }
}


__device__ __forceinline__ int32_t wave_max(int32_t v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, kWave));
  return v;
}
__device__ __forceinline__ int32_t wave_min(int32_t v) {
  for (int d = kWave / 2; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, kWave));
  return v;
}

// Clearing, compression and the per-dimension row bounds of the page schedule in one pass
// over the columns (a column's steps read only what k_lookup left and the column itself):
//  * clearing (kernel.h:483-493): flagged columns that are not pivots are zeroed;
//  * compression (kernel.h:495-520): rows of negative cells are removed.  The rows of a
//    dimension-d column are cells of dimension d-1, so only columns one dimension above a
//    dimension holding a negative cell are scanned; the others keep their rows (the removal
//    only speeds up later additions, the pairs do not depend on it);
//  * bounds: per dimension, the largest pending low + 1 and the smallest row.
__global__ void k_clear_compress_stats(Dev d, int32_t* maxrow, int32_t* minrow) {
  const uint32_t negdims = d.ctr[cw::kNegDims];
  int32_t mx[kMaxDims], mn[kMaxDims];
  for (int k = 0; k < kMaxDims; k++) {
    mx[k] = -1;
    mn[k] = 0x7fffffff;
  }
  GRID_STRIDE(c, d.ncols) {
    uint8_t st = d.state[c];
    if (st == ST_ZERO) continue;
    if (d.flag[c] && st != ST_PIVOT) {  // clearing (kernel.h:483-493)
      d.state[c] = ST_ZERO;
      d.len[c] = 0;
      d.low[c] = -1;
      continue;
    }
    int32_t* col = d.pool + d.off[c];
    int32_t l = d.len[c];
    const int32_t dm = d.dims[c];
    if (dm >= 1 && dm <= 32 && ((negdims >> (dm - 1)) & 1u)) {  // compression (kernel.h:495-520)
      int32_t o = 0;
      for (int32_t k = 0; k < l; k++) {
        const int32_t r = col[k];
        if (!d.flag[d.n + r]) col[o++] = r;
      }
      l = o;
      d.len[c] = o;
      d.low[c] = o ? col[o - 1] : -1;
      if (!o) {
        d.state[c] = ST_ZERO;
        continue;
      }
    }
    if (st == ST_PIVOT) d.pivmeta[col[l - 1]] = pivmeta_of(d.off[c], l, col);
    if (dm >= 0 && dm < kMaxDims) {
      const int32_t first = col[0], lowp1 = col[l - 1] + 1;
      for (int k = 0; k < kMaxDims; k++)
        if (k == dm) {  // static register indexing
          if (st == ST_PENDING) mx[k] = max(mx[k], lowp1);
          mn[k] = min(mn[k], first);
        }
    }
  }
  for (int k = 0; k < kMaxDims; k++) {
    const int32_t a = wave_max(mx[k]), b = wave_min(mn[k]);
    if (lane_id() == 0) {
      if (a >= 0) atomicMax(&maxrow[k], a);
      if (b != 0x7fffffff) atomicMin(&minrow[k], b);
    }
  }
}

// After a claim round: the winners publish where their columns are stored.
__global__ void k_pivmeta(Dev d, int r) {
  const int in = r & 1;
  const uint32_t cnt = d.ctr[cw::kW + in];
  const int32_t* list = d.work[in];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c = list[i];
    const int32_t l = d.low[c];
    if (d.piv[l] == (uint32_t)c) {
      const int64_t o = d.off[c];
      d.pivmeta[l] = pivmeta_of(o, d.len[c], d.pool + o);
    }
  }
}


// one atomic per wavefront (a plain per-thread atomic on one address serialises millions)
__global__ void k_maxdim(int64_t n, const int32_t* dims, int32_t* out, uint32_t* bad) {
  int32_t mx = -1, b = 0;
  // four independent loads per thread and trip: enough bytes in flight for HBM
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += 4 * stride) {
    int32_t x[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t c = base + j * stride + threadIdx.x;
      x[j] = c < n ? dims[c] : 0;
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (x[j] < 0 || x[j] >= kMaxDims) b = 1;
      else mx = max(mx, x[j]);
    }
  }
  mx = wave_max(mx);
  b = wave_max(b);
  if (lane_id() == 0) {
    if (mx >= 0) atomicMax(out, mx);
    if (b) atomicOr(bad, 1u);
  }
}


// ------------------------------------------------------------------------------------
// ordered compaction by a small key (stable partition), used for the per-dimension
// pending lists and for the pair list
// ------------------------------------------------------------------------------------
constexpr int kCompactBlock = 256;                   // threads per compaction workgroup
constexpr int kCompactSub = 4;                       // consecutive sub-tiles per workgroup
constexpr int kCompactItems = kCompactBlock * kCompactSub;

// Per workgroup and key: how many of its kCompactItems cells have that key.
template <typename KeyF>
__global__ void __launch_bounds__(kCompactBlock) k_key_count(int64_t n, KeyF key, uint32_t* counts, int nkeys) {
  __shared__ uint32_t s[kMaxDims];
  if (threadIdx.x < (unsigned)nkeys) s[threadIdx.x] = 0;
  __syncthreads();
  uint32_t mine[kMaxDims] = {};  // lane 0 of each wave accumulates its wave's counts
  for (int j = 0; j < kCompactSub; j++) {
    const int64_t i = (int64_t)blockIdx.x * kCompactItems + j * kCompactBlock + threadIdx.x;
    const int k = i < n ? key(i) : -1;
#pragma unroll
    for (int kk = 0; kk < kMaxDims; kk++)  // static indexing keeps mine[] in registers
      if (kk < nkeys) mine[kk] += (uint32_t)__popcll(__ballot(k == kk));
  }
  if (lane_id() == 0) {
#pragma unroll
    for (int kk = 0; kk < kMaxDims; kk++)
      if (kk < nkeys && mine[kk]) atomicAdd(&s[kk], mine[kk]);
  }
  __syncthreads();
  // key-major: an exclusive scan over counts[k * nb + b] is the stable partition's base of
  // (key k, workgroup b)
  if (threadIdx.x < (unsigned)nkeys) counts[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = s[threadIdx.x];
}

// Stable scatter: cells keep their order within each key (sub-tiles, then waves, then lanes).
// per-key totals from the scanned key-major counts (bases[nkeys * nb] is the grand total)
__global__ void k_key_totals(int nkeys, int64_t nb, const int64_t* bases, int64_t* totals) {
  if (threadIdx.x < (unsigned)nkeys) totals[threadIdx.x] = bases[(threadIdx.x + 1) * nb] - bases[threadIdx.x * nb];
}

template <typename KeyF, typename EmitF>
__global__ void __launch_bounds__(kCompactBlock) k_key_scatter(int64_t n, KeyF key, const int64_t* bases, int nkeys,
                                                              EmitF emit) {
  __shared__ uint32_t s_wave[kCompactBlock / kWave][kMaxDims];
  __shared__ uint32_t s_run[kMaxDims];
  const int lane = lane_id();
  const int w = threadIdx.x / kWave;
  const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (threadIdx.x < (unsigned)nkeys) s_run[threadIdx.x] = (uint32_t)bases[(int64_t)threadIdx.x * gridDim.x + blockIdx.x];
  for (int j = 0; j < kCompactSub; j++) {
    const int64_t i = (int64_t)blockIdx.x * kCompactItems + j * kCompactBlock + threadIdx.x;
    const int k = i < n ? key(i) : -1;
    uint32_t rank_in_wave = 0;
    for (int kk = 0; kk < nkeys; kk++) {
      const uint64_t m = __ballot(k == kk);
      if (k == kk) rank_in_wave = (uint32_t)__popcll(m & lower);
      if (lane == 0) s_wave[w][kk] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (k >= 0) {
      uint32_t before = s_run[k];
      for (int ww = 0; ww < w; ww++) before += s_wave[ww][k];
      emit(i, (int64_t)before + rank_in_wave);
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nkeys) {
      uint32_t t = 0;
      for (int ww = 0; ww < kCompactBlock / kWave; ww++) t += s_wave[ww][threadIdx.x];
      s_run[threadIdx.x] += t;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------
// page loop kernels
// ------------------------------------------------------------------------------------
// Active columns of the page: pending with low in the row band (specseq.h:583-607 scans
// the same window every iteration; here the window is scanned once per page).
__global__ void k_scan_window(Dev d, const int32_t* cols, int32_t ncols, int32_t minrow) {
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < ncols; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    int32_t c = -1;
    bool act = false;
    if (i < ncols) {
      c = cols[i];
      act = d.state[c] == ST_PENDING && d.low[c] >= minrow;
    }
    wave_push(act, c, d.work[0], &d.ctr[cw::kW]);
  }
}

// Claim round r: pivot claims by atomicMin on the pivot table (kernel.h:583-607).
__global__ void k_claim(Dev d, int r) {
  const int in = r & 1, nx = (r + 1) & 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    d.ctr[cw::kA + nx] = 0;
    d.ctr[cw::kW + nx] = 0;
    d.ctr[cw::kLong] = 0;  // the long and huge lists of the previous round are consumed
    d.ctr[cw::kHuge] = 0;
    d.ctr[cw::kHugeNext] = 0;
  }
  const uint32_t cnt = d.ctr[cw::kW + in];
  const int32_t* list = d.work[in];
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < cnt; base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = base + threadIdx.x;
    bool add = false;
    int32_t who = -1;
    if (i < cnt) {
      const int32_t c = list[i];
      const uint32_t old = atomicMin(&d.piv[d.low[c]], (uint32_t)c);
      if (old == kEmpty) {
        d.state[c] = ST_PIVOT;
      } else if (old > (uint32_t)c) {
        d.state[c] = ST_PIVOT;
        add = true;
        who = (int32_t)old;  // displaced pivot must now add c
      } else {
        add = true;
        who = c;
      }
    }
    wave_push(add, who, d.work[2 + in], &d.ctr[cw::kA + in]);
  }
}


// One wavefront: O = A xor B for sorted sets in HBM, merged kTile merged positions at a
// time through the wave's LDS staging buffers.  Returns |O|.  O must hold |A|+|B|.
__device__ int wave_tiled_xor(const int32_t* A, int32_t a_len, const int32_t* B, int32_t b_len, int32_t* O,
                              int32_t* bufA, int32_t* bufB, int32_t* bufO) {
  int32_t ia = 0, jb = 0, written = 0;
  while (ia < a_len || jb < b_len) {
    const int32_t ca = min(kTile, a_len - ia), cb = min(kTile, b_len - jb);
    wave_copy(bufA, A + ia, ca);
    wave_copy(bufB, B + jb, cb);
    wave_sync();
    const int32_t K = min(kTile, ca + cb);
    int32_t jE;
    const int32_t iE = split_adj(bufA, ca, bufB, cb, K, &jE);
    const int32_t m = wave_xor_lds(bufA, iE, bufB, jE, bufO);
    wave_copy(O + written, bufO, m);
    wave_sync();
    written += m;
    ia += iE;
    jb += jE;
  }
  __threadfence();
  return written;
}

// Add round r: one wavefront per listed column; it adds piv[low] while that pivot is a
// smaller column and its low stays in the row band (kernel.h:814-1065 does one addition per
// warp, or a block-wide chase).  The running column stays in LDS between additions while it
// and the pivot fit; longer ones merge tile by tile into a fresh store slot.
template <int kAddWaves>
__global__ void __launch_bounds__(kWave * kAddWaves) k_add(Dev d, int r, int32_t minrow) {
  __shared__ int32_t lds[kAddWaves][3][kTileBuf];
  const int in = r & 1, nx = (r + 1) & 1;
  const int w = threadIdx.x / kWave;
  const int lane = lane_id();
  const uint32_t cnt = d.ctr[cw::kA + in];
  const int32_t* list = d.work[2 + in];
  const int64_t gw = (int64_t)blockIdx.x * kAddWaves + w;
  const int64_t nw = (int64_t)gridDim.x * kAddWaves;
  int32_t* bufA = lds[w][0];
  int32_t* bufB = lds[w][1];
  int32_t* bufO = lds[w][2];
  unsigned long long n_add = 0, n_bytes = 0;
  uint32_t max_chain = 0, max_len = 0;

  for (int64_t it = gw; it < cnt; it += nw) {
    const int32_t c = list[it];
    int64_t a_off = d.off[c];
    int32_t a_cap = d.cap[c];
    int32_t a_len = d.len[c];
    int32_t l = d.low[c];
    bool in_lds = false;
    bool failed = false;
    int32_t* sA = bufA;
    int32_t* sO = bufO;
    uint32_t chain = 0;
    bool deferred = false;
    // a2: the running column's largest row below l (-1 none).  When it differs from the
    // pivot's, the larger of the two is the next low: that low's pivot is fetched together
    // with the current pivot's rows.
    int32_t a2 = a_len >= 2 ? d.pool[a_off + a_len - 2] : -1;
    int32_t spec_row = -2;
    uint32_t spec_p = kEmpty;
    PivMeta spec_pm{};
    for (;;) {
      uint32_t p;
      PivMeta pm;
      if (l == spec_row) {
        p = spec_p;
        pm = spec_pm;
      } else {
        p = d.piv[l];
        pm = d.pivmeta[l];
      }
      if (p == kEmpty || p >= (uint32_t)c) break;
      const int32_t b_len = pm.len;
      const int32_t* B = d.pool + pm.off;
      spec_row = -2;
      if (a2 != pm.second) {
        const int32_t cand = max(a2, pm.second);
        if (cand >= 0 && cand >= minrow) {
          spec_row = cand;
          spec_p = d.piv[cand];
          spec_pm = d.pivmeta[cand];
        }
      }
      int32_t r_len;
      if (a_len + b_len <= d.short_max) {
        if (!in_lds) {
          wave_copy(sA, d.pool + a_off, a_len);
          in_lds = true;
        }
        wave_copy(bufB, B, b_len);
        wave_sync();
        r_len = wave_xor_lds(sA, a_len, bufB, b_len, sO);
        int32_t* t = sA;
        sA = sO;
        sO = t;
      } else {
        // too long for this wave's LDS: the column moves to the long list, where a whole
        // workgroup continues the chase in the same add round
        if (in_lds) {
          if (a_len > a_cap && !wave_alloc(d, a_len, &a_off, &a_cap)) { failed = true; break; }
          wave_copy(d.pool + a_off, sA, a_len);
          in_lds = false;
          __threadfence();
        }
        deferred = true;
        break;
      }
      n_add += 1;
      chain += 1;
      max_len = max(max_len, (uint32_t)(a_len + b_len));
      n_bytes += 4ull * (unsigned long long)(a_len + b_len + r_len) + 24ull;
      a_len = r_len;
      if (a_len == 0) {
        l = -1;
        break;
      }
      l = sA[a_len - 1];
      a2 = a_len >= 2 ? sA[a_len - 2] : -1;
      if (l < minrow) break;
    }
    max_chain = max(max_chain, chain);
    if (!failed && in_lds) {
      if (a_len > a_cap && !wave_alloc(d, a_len, &a_off, &a_cap)) failed = true;
      else wave_copy(d.pool + a_off, sA, a_len);
    }
    if (failed) {  // store exhausted: the call reports SSQ_ENOMEM; keep memory safe
      if (lane == 0) {
        d.len[c] = 0;
        d.low[c] = -1;
        d.state[c] = ST_ZERO;
      }
      continue;
    }
    if (deferred) {
      if (lane == 0) {
        d.off[c] = a_off;
        d.cap[c] = a_cap;
        d.len[c] = a_len;
        d.low[c] = l;
        d.state[c] = ST_PENDING;
        const uint32_t pos = atomicAdd(&d.ctr[cw::kLong], 1u);
        d.work[4][pos] = c;
      }
      wave_sync();
      continue;
    }
    bool again = false;
    if (lane == 0) {
      d.off[c] = a_off;
      d.cap[c] = a_cap;
      d.len[c] = a_len;
      d.low[c] = a_len ? l : -1;
      d.state[c] = a_len ? ST_PENDING : ST_ZERO;
      if (a_len && l >= minrow) {
        const uint32_t p = d.piv[l];
        again = (p == kEmpty || p > (uint32_t)c);
      }
    }
    if (lane == 0 && again) {
      const uint32_t pos = atomicAdd(&d.ctr[cw::kW + nx], 1u);
      d.work[nx][pos] = c;
    }
    wave_sync();
  }
  if (lane == 0 && n_add) {
    atomicAdd(&d.stats[0], n_add);
    atomicAdd(&d.stats[1], n_bytes);
    atomicMax(&d.ctr[cw::kChain], max_chain);
    atomicMax(&d.ctr[cw::kChain + 1], max_len);
  }
}


// Long columns of an add round: one workgroup per column.  While the column's rows lie in a
// window of kLongBits rows below its low, it is kept as a bit set in LDS: adding a pivot
// toggles the pivot's rows (LDS atomics) and the new low is the highest set bit, found by a
// block-wide scan downwards from the old low.  This reads only the pivot per addition
// (a sorted merge reads and writes the whole running column).  A pivot with rows below the
// window sends the column back to sorted form and tiled merges.
constexpr int kLongBits = 1 << 19;
constexpr int kLongWords = kLongBits / 32;
constexpr int kLongThreads = 512;
constexpr int kHugeGroups = 256;        // k_add_huge workgroups, one device bit set region each
constexpr int32_t kHugeBits = 1 << 27;  // rows of one region (16 MiB)

__device__ __forceinline__ int32_t block_max_i32(int32_t v, int32_t* scratch) {
  v = wave_max(v);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane_id() == 0) scratch[w] = v;
  __syncthreads();
  int32_t m = scratch[0];
  for (int k = 1; k < kLongThreads / kWave; k++) m = max(m, scratch[k]);
  return m;
}

// Top two of distinct values (-1 = none): (a1 > a2) merged with (b1 > b2).
__device__ __forceinline__ void top2_merge(int32_t& a1, int32_t& a2, int32_t b1, int32_t b2) {
  a2 = max(min(a1, b1), max(a2, b2));
  a1 = max(a1, b1);
}
__device__ __forceinline__ void block_top2(int32_t& v1, int32_t& v2, int32_t* scratch) {
  for (int k = kWave / 2; k > 0; k >>= 1) {
    const int32_t o1 = __shfl_xor(v1, k, kWave), o2 = __shfl_xor(v2, k, kWave);
    top2_merge(v1, v2, o1, o2);
  }
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane_id() == 0) {
    scratch[2 * w] = v1;
    scratch[2 * w + 1] = v2;
  }
  __syncthreads();
  v1 = scratch[0];
  v2 = scratch[1];
  for (int k = 1; k < kLongThreads / kWave; k++) top2_merge(v1, v2, scratch[2 * k], scratch[2 * k + 1]);
}

__device__ __forceinline__ int32_t block_excl_scan(int32_t v, int32_t* scratch, int32_t* total) {
  const int w = threadIdx.x / kWave;
  const int incl = wave_incl_scan(v);
  __syncthreads();
  if (lane_id() == kWave - 1) scratch[w] = incl;
  __syncthreads();
  int32_t base = 0, tot = 0;
  for (int k = 0; k < kLongThreads / kWave; k++) {
    if (k < w) base += scratch[k];
    tot += scratch[k];
  }
  *total = tot;
  return base + incl - v;
}

__global__ void __launch_bounds__(kLongThreads) k_add_long(Dev d, int r, int32_t minrow) {
  __shared__ uint32_t bits[kLongWords];
  __shared__ int32_t red[2 * (kLongThreads / kWave)];
  __shared__ int32_t sh_count, sh_flag;
  __shared__ int64_t sh_off;
  __shared__ int32_t sh_cap;
  const int nx = (r + 1) & 1;
  const int tid = threadIdx.x;
  const uint32_t cnt = d.ctr[cw::kLong];
  const int32_t* list = d.work[4];
  unsigned long long n_add = 0, n_dense = 0, n_dbytes = 0, n_tiled = 0;
  uint32_t max_chain = 0, max_len = 0;

  for (int64_t it = blockIdx.x; it < cnt; it += gridDim.x) {
    const int32_t c = list[it];
    int64_t a_off = d.off[c];
    int32_t a_cap = d.cap[c];
    int32_t a_len = d.len[c];
    int32_t l = d.low[c];
    bool to_huge = false;
    const int32_t lo = l - d.long_window + 1;  // window [lo, l] of the bit set
    bool dense = d.pool[a_off] >= lo;
    bool failed = false;
    uint32_t chain = 0;
    // a2: the running column's largest row below l (-1 none, kUnknown not known).  When it
    // differs from the pivot's, the larger of the two is the next low, so that low's pivot
    // is fetched while the current addition is still toggling rows.
    constexpr int32_t kUnknown = -3;
    int32_t a2 = a_len >= 2 ? d.pool[a_off + a_len - 2] : -1;
    int32_t spec_row = -2;
    uint32_t spec_p = kEmpty;
    PivMeta spec_pm{};
    if (dense) {
      const int32_t nw = ((l - lo) >> 5) + 1;
      for (int32_t w = tid; w < nw; w += kLongThreads) bits[w] = 0u;
      __syncthreads();
      const int32_t* A = d.pool + a_off;
      for (int32_t k = tid; k < a_len; k += kLongThreads) {
        const int32_t x = A[k] - lo;
        atomicOr(&bits[x >> 5], 1u << (x & 31));
      }
      if (tid == 0) sh_count = a_len;
      __syncthreads();
    }
    for (;;) {
      // the chase state is uniform across the workgroup: scalar loads
      l = __builtin_amdgcn_readfirstlane(l);
      uint32_t p;
      PivMeta pm;
      if (l == spec_row) {
        p = spec_p;
        pm = spec_pm;
      } else {
        p = d.piv[l];
        pm = d.pivmeta[l];
      }
      p = (uint32_t)__builtin_amdgcn_readfirstlane((int)p);
      if (p == kEmpty || p >= (uint32_t)c) break;
      const int32_t b_len = __builtin_amdgcn_readfirstlane(pm.len);
      const int32_t b_first = __builtin_amdgcn_readfirstlane(pm.first);
      const int32_t b2 = __builtin_amdgcn_readfirstlane(pm.second);
      const int64_t b_off = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(pm.off & 0xffffffff)) |
                            ((int64_t)__builtin_amdgcn_readfirstlane((int)(pm.off >> 32)) << 32);
      const int32_t* B = d.pool + b_off;
      if (dense && b_first < lo) {
        // leave the bit set: write the column back in sorted form
        int32_t W = ((l - lo) >> 5) + 1, total;
        const int32_t per = (W + kLongThreads - 1) / kLongThreads;
        const int32_t w0 = min(W, tid * per), w1 = min(W, w0 + per);
        int32_t pc = 0;
        for (int32_t w = w0; w < w1; w++) pc += __popc(bits[w]);
        const int32_t pos0 = block_excl_scan(pc, red, &total);
        if (tid == 0) {
          int64_t o = a_off;
          int32_t cp = a_cap;
          sh_flag = 1;
          if (total > cp && !store_alloc(d, total, &o, &cp)) sh_flag = 0;
          sh_off = o;
          sh_cap = cp;
        }
        __syncthreads();
        if (!sh_flag) { failed = true; break; }
        a_off = sh_off;
        a_cap = sh_cap;
        int32_t* O = d.pool + a_off;
        int32_t pos = pos0;
        for (int32_t w = w0; w < w1; w++) {
          uint32_t x = bits[w];
          while (x) {
            const int b = __ffs(x) - 1;
            O[pos++] = lo + (w << 5) + b;
            x &= x - 1;
          }
        }
        __threadfence();
        __syncthreads();
        a_len = total;
        dense = false;
      }
      int32_t r_len;
      if (dense) {
        spec_row = -2;
        if (a2 != kUnknown && a2 != b2) {
          const int32_t cand = max(a2, b2);
          if (cand >= 0 && cand >= minrow) {
            spec_row = cand;
            spec_p = d.piv[cand];
            spec_pm = d.pivmeta[cand];
          }
        }
        // toggle the pivot's rows
        int32_t delta = 0;
        constexpr int kU = 16;  // independent loads in flight per thread
        for (int32_t k0 = tid; k0 < b_len; k0 += kLongThreads * kU) {
          int32_t x[kU];
#pragma unroll
          for (int u = 0; u < kU; u++) {
            const int32_t k = k0 + u * kLongThreads;
            x[u] = k < b_len ? B[k] - lo : -1;
          }
#pragma unroll
          for (int u = 0; u < kU; u++)
            if (x[u] >= 0) {
              const uint32_t m = 1u << (x[u] & 31);
              const uint32_t old = atomicXor(&bits[x[u] >> 5], m);
              delta += (old & m) ? -1 : 1;
            }
        }
        {  // the wavefront's sum of the changes by bit planes: |delta| <= 1024 (a pivot in the
           // bit set has at most kLongBits rows over kLongThreads threads), offset to >= 0
          const uint32_t u = (uint32_t)(delta + 2048);
          uint32_t sum = 0;
#pragma unroll
          for (int b = 0; b < 12; b++) sum += (uint32_t)__popcll(__ballot((u >> b) & 1u)) << b;
          if (lane_id() == 0) atomicAdd(&sh_count, (int32_t)sum - 2048 * kWave);
        }
        __syncthreads();
        r_len = sh_count;
        // new low and the row below it: the two highest set bits, scanning down from the old
        // low's word.  Wavefront 0 alone scans, 64 words per step (one ballot finds the
        // nonzero words), and publishes them; the second is looked for up to 8 steps past the
        // first.  Two barriers per addition:
        // every read of this step's LDS values happens before the second, every write of the
        // next step after it.
        if (tid < kWave) {
          const int lane = tid;
          int32_t t1 = -1, t2 = -1, past = 0;
          int32_t base = r_len > 0 ? (l - lo) >> 5 : -1;  // r_len is uniform (read from LDS)
          for (; base >= 0; base -= kWave) {  // t1, t2 are uniform in the wavefront: so are the exits
            // lane k holds word base - k, so rows descend with the lane: the highest rows are
            // in the first lanes whose words are nonzero, read one word at a time (scalar)
            const int32_t w = base - lane;
            const uint32_t m = w >= 0 ? bits[w] : 0u;
            uint64_t nz = __ballot(m != 0);
            while (nz && t2 < 0) {
              const int fl = __ffsll((unsigned long long)nz) - 1;
              nz &= nz - 1;
              uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)m, fl);
              const int32_t row0 = lo + ((base - fl) << 5);
              while (mm && t2 < 0) {
                const int h = 31 - __clz(mm);
                mm &= ~(1u << h);
                if (t1 < 0) t1 = row0 + h;
                else t2 = row0 + h;
              }
            }
            if (t2 >= 0 || (t1 >= 0 && r_len == 1)) break;
            if (t1 >= 0 && ++past >= 8) break;
          }
          if (lane == 0) {
            red[0] = t1;
            red[1] = t2 >= 0 ? t2 : ((base < 0 || r_len <= 1) ? -1 : kUnknown);
          }
        }
        __syncthreads();
        const int32_t newl = red[0];
        a2 = red[1];
        n_dense += 1;
        n_dbytes += 4ull * (unsigned long long)b_len + 24ull;  // the pivot read; the bit set is in LDS
        max_len = max(max_len, (uint32_t)(a_len + b_len));
        chain += 1;
        a_len = r_len;
        l = r_len ? newl : -1;
      } else {
        // the column's rows leave the window: k_add_huge continues the chase in this round
        to_huge = true;
        break;
      }
      if (l < 0 || l < minrow) break;
    }
    max_chain = max(max_chain, chain);
    if (!failed && dense) {  // write the bit set back in sorted form
      const int32_t W = l >= 0 ? ((l - lo) >> 5) + 1 : 0;
      int32_t total;
      const int32_t per = (W + kLongThreads - 1) / kLongThreads;
      const int32_t w0 = min(W, tid * per), w1 = min(W, w0 + per);
      int32_t pc = 0;
      for (int32_t w = w0; w < w1; w++) pc += __popc(bits[w]);
      const int32_t pos0 = block_excl_scan(pc, red, &total);
      if (tid == 0) {
        int64_t o = a_off;
        int32_t cp = a_cap;
        sh_flag = 1;
        if (total > cp && !store_alloc(d, total, &o, &cp)) sh_flag = 0;
        sh_off = o;
        sh_cap = cp;
      }
      __syncthreads();
      if (!sh_flag) failed = true;
      else {
        a_off = sh_off;
        a_cap = sh_cap;
        int32_t* O = d.pool + a_off;
        int32_t pos = pos0;
        for (int32_t w = w0; w < w1; w++) {
          uint32_t x = bits[w];
          while (x) {
            const int b = __ffs(x) - 1;
            O[pos++] = lo + (w << 5) + b;
            x &= x - 1;
          }
        }
        a_len = total;
      }
    }
    if (tid == 0) {
      if (failed) {
        d.len[c] = 0;
        d.low[c] = -1;
        d.state[c] = ST_ZERO;
      } else {
        d.off[c] = a_off;
        d.cap[c] = a_cap;
        d.len[c] = a_len;
        d.low[c] = a_len ? l : -1;
        d.state[c] = a_len ? ST_PENDING : ST_ZERO;
        if (to_huge) {
          const uint32_t pos = atomicAdd(&d.ctr[cw::kHuge], 1u);
          d.work[5][pos] = c;
        } else if (a_len && l >= minrow) {
          const uint32_t p = d.piv[l];
          if (p == kEmpty || p > (uint32_t)c) {
            const uint32_t pos = atomicAdd(&d.ctr[cw::kW + nx], 1u);
            d.work[nx][pos] = c;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0 && (n_add || n_dense)) {
    atomicAdd(&d.stats[0], n_add + n_dense);
    atomicAdd(&d.stats[2], n_tiled);
    atomicAdd(&d.stats[3], n_dense);
    atomicAdd(&d.stats[4], n_dbytes);
    atomicMax(&d.ctr[cw::kChain], max_chain);
    atomicMax(&d.ctr[cw::kChain + 1], max_len);
  }
}

// Columns whose rows spread past the LDS window (kLongBits rows): one workgroup per column,
// taken one at a time (their chases differ in length by orders of magnitude).  The running
// column is a bit set over a region of gspan rows whose top row is the column's low.  The top
// kLongBits rows of it (the window) live in LDS; the rows below, the tail, live in this
// workgroup's region of device memory, with a summary in LDS of which 32-word chunks may be
// nonzero.  An addition toggles the pivot's rows: LDS atomics in the window, global atomics
// below it that nothing waits for.  The new low is looked for in the window.  Only when the
// window has emptied does the workgroup wait for its global toggles, find the tail's highest
// row through the summary, and move the window down onto it (reading those words into LDS and
// clearing them).  The next pivot is fetched while the current one is toggled, whenever the
// column's second row is known (the next low is then the larger of the two second rows when
// they differ).  A region is all zero between columns.  A pivot with rows below the region
// sends the column to sorted form and tiled merges, ping-ponging between the column's own slot
// and one scratch slot.
__device__ __forceinline__ uint32_t gload(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // past the L1, after atomics
}
__device__ __forceinline__ void gxor(uint32_t* p, uint32_t m) {  // no result: nothing waits for it
  (void)__hip_atomic_fetch_xor(p, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Summary of the tail in LDS: bit j of summ set <=> region words [32j, 32j + 32) (rows 1024j ..
// 1024j + 1023 above the region's base) may hold set bits; a clear bit means those words are
// zero.  Every global toggle marks its chunk; searches clear the stale bits they meet.
constexpr int kChunkWords = 32;
constexpr int kSumWords = (kHugeBits / 32 / kChunkWords + 31) / 32;  // 4096 words (16 KiB)
__device__ __forceinline__ void mark_chunk(uint32_t* summ, int32_t g) {
  atomicOr(&summ[g >> 10], 1u << ((g >> 5) & 31));
}
__device__ __forceinline__ bool chunk_marked(const uint32_t* summ, int32_t ch) {
  return (summ[ch >> 5] >> (ch & 31)) & 1u;
}

// Wavefront 0: the two highest set bits of the window at or below LDS word top_w (rows
// wrow0 + 32 * word + bit).  t2 is kUnknown when the window's bottom was reached with t1 only
// (the tail may hold it) or when it lies more than 8 steps of 64 words below t1.
constexpr int32_t kUnknown = -3;
__device__ __forceinline__ void window_top2(const uint32_t* bits, int32_t top_w, int32_t wrow0, int32_t* t1o,
                                            int32_t* t2o) {
  const int lane = lane_id();
  int32_t t1 = -1, t2 = -1, past = 0;
  int32_t base = top_w;
  for (; base >= 0; base -= kWave) {  // t1, t2 are uniform in the wavefront: so are the exits
    const int32_t w = base - lane;    // lane k holds word base - k: rows descend with the lane
    const uint32_t m = w >= 0 ? bits[w] : 0u;
    uint64_t nz = __ballot(m != 0);
    while (nz && t2 < 0) {
      const int fl = __ffsll((unsigned long long)nz) - 1;
      nz &= nz - 1;
      uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)m, fl);
      const int32_t row0 = wrow0 + ((base - fl) << 5);
      while (mm && t2 < 0) {
        const int h = 31 - __clz(mm);
        mm &= ~(1u << h);
        if (t1 < 0) t1 = row0 + h;
        else t2 = row0 + h;
      }
    }
    if (t2 >= 0) break;
    if (t1 >= 0 && ++past >= 8) break;
  }
  *t1o = t1;
  *t2o = t2 >= 0 ? t2 : (t1 >= 0 ? kUnknown : -1);
}

// Wavefront 0: the tail's highest set row below region chunk ch_top + 1 (all global toggles
// made visible beforehand), -1 when the tail is empty.  Chunks found empty are unmarked.
__device__ __forceinline__ int32_t tail_top(uint32_t* summ, const uint32_t* gb, int32_t ch_top, int32_t nwords,
                                            int32_t lo) {
  const int lane = lane_id();
  int32_t ch = ch_top;
  while (ch >= 0) {
    int32_t found = -1;
    for (int32_t sb = ch >> 5; sb >= 0 && found < 0; sb -= kWave) {
      const int32_t sw = sb - lane;
      uint32_t m = sw >= 0 ? summ[sw] : 0u;
      if (sw == (ch >> 5) && (ch & 31) != 31) m &= (2u << (ch & 31)) - 1u;
      const uint64_t nz = __ballot(m != 0u);
      if (nz) {
        const int fl = __ffsll((unsigned long long)nz) - 1;
        const uint32_t mm = (uint32_t)__builtin_amdgcn_readlane((int)m, fl);
        found = ((sb - fl) << 5) + (31 - __clz(mm));
      }
    }
    if (found < 0) return -1;
    const int32_t w0 = found << 5;
    const uint32_t x = (lane < kChunkWords && w0 + lane < nwords) ? gload(gb + w0 + lane) : 0u;
    const uint64_t nzw = __ballot(x != 0u);
    if (nzw) {
      const int hi = 63 - __clzll((long long)nzw);
      const uint32_t xx = (uint32_t)__builtin_amdgcn_readlane((int)x, hi);
      return lo + ((w0 + hi) << 5) + (31 - __clz(xx));
    }
    if (lane == 0) summ[found >> 5] &= ~(1u << (found & 31));
    ch = found - 1;
  }
  return -1;
}

// The column back in sorted form: the tail's marked chunks of region words [w_lo, ww0) (each
// word read is cleared), then the window's words up to LDS word top_w.  Returns false when the
// store is exhausted (the region is cleared all the same).
__device__ bool huge_write_back(Dev& d, uint32_t* gb, const uint32_t* summ, const uint32_t* bits, int32_t lo,
                                int32_t ww0, int32_t w_lo, int32_t top_w, int64_t* a_off, int32_t* a_cap,
                                int32_t* a_len, int32_t* red, int32_t* sh_flag, int64_t* sh_off, int32_t* sh_cap) {
  const int tid = threadIdx.x;
  const int32_t nwords = d.gspan >> 5;
  __threadfence();  // the last addition's global toggles, before the tail is read
  __syncthreads();
  // tail chunks [c_lo, c_hi] (ww0 is chunk-aligned: the tail ends at a chunk boundary)
  const int32_t c_lo = max(w_lo, 0) >> 5, c_hi = (ww0 >> 5) - 1;
  const int32_t C = c_hi >= c_lo ? c_hi - c_lo + 1 : 0;
  const int32_t per = (C + kLongThreads - 1) / kLongThreads;
  const int32_t v0 = c_lo + min(C, tid * per), v1 = c_lo + min(C, tid * per + per);
  const int32_t WW = top_w >= 0 ? top_w + 1 : 0;  // window words in use
  const int32_t wper = (WW + kLongThreads - 1) / kLongThreads;
  const int32_t u0 = min(WW, tid * wper), u1 = min(WW, tid * wper + wper);
  int32_t pt = 0, pw = 0;
  for (int32_t ch = v0; ch < v1; ch++) {
    if (!chunk_marked(summ, ch)) continue;
    const int32_t ke = min(kChunkWords, nwords - (ch << 5));
    for (int32_t k = 0; k < ke; k++) pt += __popc(gload(gb + (ch << 5) + k));
  }
  for (int32_t w = u0; w < u1; w++) pw += __popc(bits[w]);
  int32_t total_t, total_w;
  const int32_t post = block_excl_scan(pt, red, &total_t);
  const int32_t posw = block_excl_scan(pw, red, &total_w);
  const int32_t total = total_t + total_w;
  if (tid == 0) {
    int64_t o = *a_off;
    int32_t cp = *a_cap;
    *sh_flag = 1;
    if (total > cp && !store_alloc(d, total, &o, &cp)) *sh_flag = 0;
    *sh_off = o;
    *sh_cap = cp;
  }
  __syncthreads();
  const bool ok = *sh_flag != 0;
  int32_t* O = d.pool + *sh_off;
  int32_t pos = post;
  for (int32_t ch = v0; ch < v1; ch++) {
    if (!chunk_marked(summ, ch)) continue;
    const int32_t ke = min(kChunkWords, nwords - (ch << 5));
    for (int32_t k = 0; k < ke; k++) {
      const int32_t g = (ch << 5) + k;
      uint32_t x = gload(gb + g);
      if (!x) continue;
      gb[g] = 0u;
      if (ok)
        while (x) {
          const int b = __ffs(x) - 1;
          O[pos++] = lo + (g << 5) + b;
          x &= x - 1;
        }
    }
  }
  pos = total_t + posw;
  const int32_t wrow0 = lo + (ww0 << 5);
  if (ok)
    for (int32_t w = u0; w < u1; w++) {
      uint32_t x = bits[w];
      while (x) {
        const int b = __ffs(x) - 1;
        O[pos++] = wrow0 + (w << 5) + b;
        x &= x - 1;
      }
    }
  __threadfence();
  __syncthreads();
  if (ok) {
    *a_off = *sh_off;
    *a_cap = *sh_cap;
    *a_len = total;
  }
  return ok;
}

// the window's first region word for a low in region word lw: lw's chunk is the window's top
__device__ __forceinline__ int32_t window_base(int32_t lw) {
  return max(0, ((lw >> 5) + 1) * kChunkWords - kLongWords);
}

__global__ void __launch_bounds__(kLongThreads) k_add_huge(Dev d, int r, int32_t minrow) {
  __shared__ uint32_t bits[kLongWords];
  __shared__ uint32_t summ[kSumWords];
  __shared__ int32_t stage[3][kTileBuf];
  __shared__ int32_t red[2 * (kLongThreads / kWave)];
  __shared__ int32_t sh_flag, sh_minw, sh_cap;
  __shared__ int64_t sh_off;
  __shared__ uint32_t sh_item;
  const int nx = (r + 1) & 1;
  const int tid = threadIdx.x;
  const uint32_t cnt = d.ctr[cw::kHuge];
  const int32_t* list = d.work[5];
  uint32_t* gb = d.gbits + (int64_t)blockIdx.x * (d.gspan >> 5);
  const int32_t nwords = d.gspan >> 5;
  const int32_t nsum = ((nwords + kChunkWords - 1) / kChunkWords + 31) >> 5;
  unsigned long long n_add = 0, n_tiled = 0, n_bytes = 0;
  uint32_t max_chain = 0, max_len = 0;
  for (;;) {
    if (tid == 0) sh_item = atomicAdd(&d.ctr[cw::kHugeNext], 1u);
    __syncthreads();
    const uint32_t it = sh_item;
    if (it >= cnt) break;
    const int32_t c = list[it];
    int64_t a_off = d.off[c];
    int32_t a_cap = d.cap[c];
    int32_t a_len = d.len[c];
    int32_t l = d.low[c];
    const int32_t lo = l - d.gspan + 1;  // the region holds rows [lo, l]
    bool dense = a_len > 0 && d.pool[a_off] >= lo;
    bool failed = false;
    uint32_t chain = 0;
    int64_t t_off = 0;  // scratch slot of the tiled merges (thread 0's copy is used)
    int32_t t_cap = 0;
    int32_t ww0 = window_base((l - lo) >> 5);  // region word of LDS word 0
    int32_t a2 = kUnknown;                     // the running column's second row
    int32_t spec_row = -2;
    uint32_t spec_p = kEmpty;
    PivMeta spec_pm{};
    if (dense) {
      for (int32_t k = tid; k < kLongWords; k += kLongThreads) bits[k] = 0u;
      for (int32_t k = tid; k < nsum; k += kLongThreads) summ[k] = 0u;
      __syncthreads();
      const int32_t* A = d.pool + a_off;
      for (int32_t k = tid; k < a_len; k += kLongThreads) {
        const int32_t x = A[k] - lo, g = x >> 5;
        if (g >= ww0) {
          atomicOr(&bits[g - ww0], 1u << (x & 31));
        } else {
          atomicOr(gb + g, 1u << (x & 31));
          mark_chunk(summ, g);
        }
      }
      if (tid == 0) sh_minw = (A[0] - lo) >> 5;
      __threadfence();
      __syncthreads();
      a2 = a_len >= 2 ? A[a_len - 2] : -1;
    }
    for (;;) {
      l = __builtin_amdgcn_readfirstlane(l);
      uint32_t p;
      PivMeta pm;
      if (l == spec_row) {
        p = spec_p;
        pm = spec_pm;
      } else {
        p = d.piv[l];
        pm = d.pivmeta[l];  // fetched with the pivot (meaningful only if it exists)
      }
      p = (uint32_t)__builtin_amdgcn_readfirstlane((int)p);
      if (p == kEmpty || p >= (uint32_t)c) break;
      const int32_t b_len = __builtin_amdgcn_readfirstlane(pm.len);
      const int32_t b_first = __builtin_amdgcn_readfirstlane(pm.first);
      const int32_t b2 = __builtin_amdgcn_readfirstlane(pm.second);
      const int64_t b_off = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(pm.off & 0xffffffff)) |
                            ((int64_t)__builtin_amdgcn_readfirstlane((int)(pm.off >> 32)) << 32);
      const int32_t* B = d.pool + b_off;
      if (dense && b_first < lo) {  // leave the region: sorted form, then tiled merges
        if (!huge_write_back(d, gb, summ, bits, lo, ww0, sh_minw, ((l - lo) >> 5) - ww0, &a_off, &a_cap, &a_len,
                             red, &sh_flag, &sh_off, &sh_cap)) {
          failed = true;
          break;
        }
        dense = false;
      }
      if (dense) {
        spec_row = -2;
        if (a2 != kUnknown && a2 != b2) {
          const int32_t cand = max(a2, b2);
          if (cand >= 0 && cand >= minrow) {
            spec_row = cand;
            spec_p = d.piv[cand];
            spec_pm = d.pivmeta[cand];
          }
        }
        // toggle the pivot's rows
        constexpr int kU = 16;  // independent loads in flight per thread
        for (int32_t k0 = tid; k0 < b_len; k0 += kLongThreads * kU) {
          int32_t x[kU];
#pragma unroll
          for (int u = 0; u < kU; u++) {
            const int32_t k = k0 + u * kLongThreads;
            x[u] = k < b_len ? B[k] - lo : -1;
          }
#pragma unroll
          for (int u = 0; u < kU; u++)
            if (x[u] >= 0) {
              const int32_t g = x[u] >> 5;
              const uint32_t m = 1u << (x[u] & 31);
              if (g >= ww0) {
                atomicXor(&bits[g - ww0], m);
              } else {
                gxor(gb + g, m);
                mark_chunk(summ, g);
              }
            }
        }
        if (tid == 0) atomicMin(&sh_minw, (b_first - lo) >> 5);
        __syncthreads();
        // new low and the row below it, in the window; consecutive additions use alternate
        // slots of red[], so an addition's slot is rewritten only two barriers after every
        // thread has read it
        const int rb = (int)(chain & 1u) * 2;
        if (tid < kWave) {
          int32_t t1, t2;
          window_top2(bits, ((l - lo) >> 5) - ww0, lo + (ww0 << 5), &t1, &t2);
          if (tid == 0) {
            red[rb] = t1;
            red[rb + 1] = t2;
          }
        }
        __syncthreads();
        int32_t newl = red[rb];
        a2 = red[rb + 1];
        if (newl < 0 && ww0 > 0) {
          // the window is empty: wait for the global toggles, take the tail's highest row and
          // move the window onto it
          __threadfence();
          __syncthreads();
          if (tid < kWave) {
            const int32_t t = tail_top(summ, gb, (ww0 >> 5) - 1, nwords, lo);
            if (tid == 0) red[0] = t;
          }
          __syncthreads();
          newl = red[0];
          if (newl >= 0) {
            const int32_t lw = (newl - lo) >> 5;
            const int32_t nb = window_base(lw);  // chunk-aligned, below the old window
            // pull the marked chunks [nb >> 5, lw >> 5] into the (all zero) window
            const int32_t c0 = nb >> 5, C = (lw >> 5) - c0 + 1;
            for (int32_t ci = tid; ci < C; ci += kLongThreads) {
              const int32_t ch = c0 + ci;
              if (!chunk_marked(summ, ch)) continue;
              const int32_t ke = min(kChunkWords, nwords - (ch << 5));
              for (int32_t k = 0; k < ke; k++) {
                const int32_t g = (ch << 5) + k;
                const uint32_t x = gload(gb + g);
                bits[g - nb] = x;
                if (x) gb[g] = 0u;
              }
            }
            __syncthreads();
            for (int32_t ci = tid; ci < C; ci += kLongThreads) {  // every pulled chunk is now zero
              const int32_t ch = c0 + ci;
              if (chunk_marked(summ, ch)) atomicAnd(&summ[ch >> 5], ~(1u << (ch & 31)));
            }
            ww0 = nb;
            __syncthreads();
            if (tid < kWave) {
              int32_t t1, t2;
              window_top2(bits, lw - ww0, lo + (ww0 << 5), &t1, &t2);
              if (tid == 0) {
                red[0] = t1;
                red[1] = t2;
              }
            }
            __syncthreads();
            newl = red[0];
            a2 = red[1];
          }
        }
        n_add += 1;
        n_bytes += 4ull * (unsigned long long)b_len + 24ull;  // the pivot read; the bit set stays on chip
        max_len = max(max_len, (uint32_t)b_len);
        chain += 1;
        l = newl;
      } else {
        // tiled merge by wave 0 into the other of two slots: the column's own and one scratch
        // slot, swapped after each addition; the scratch slot is replaced (twice the need)
        // only when a merge would not fit
        if (tid == 0) {
          sh_flag = 1;
          const int32_t need = a_len + b_len;
          if (need > t_cap) {
            int64_t o;
            int32_t cp;
            const int32_t want = need <= 0x3fffffff ? 2 * need : need;
            sh_flag = store_alloc(d, want, &o, &cp) ? 1 : 0;
            t_off = o;
            t_cap = cp;
          }
          sh_off = t_off;
          sh_cap = t_cap;
        }
        __syncthreads();
        if (!sh_flag) {
          failed = true;
          break;
        }
        int32_t r_len = 0;
        if (tid < kWave)
          r_len = wave_tiled_xor(d.pool + a_off, a_len, B, b_len, d.pool + sh_off, stage[0], stage[1], stage[2]);
        if (tid == 0) {
          red[0] = r_len;
          t_off = a_off;  // the column's previous storage becomes the scratch slot
          t_cap = a_cap;
        }
        __syncthreads();
        r_len = red[0];
        n_tiled += 1;
        n_add += 1;
        n_bytes += 4ull * (unsigned long long)(a_len + b_len + r_len) + 24ull;
        max_len = max(max_len, (uint32_t)(a_len + b_len));
        chain += 1;
        a_off = sh_off;
        a_cap = sh_cap;
        a_len = r_len;
        l = r_len ? d.pool[a_off + r_len - 1] : -1;
        __syncthreads();
      }
      if (l < 0 || l < minrow) break;
    }
    max_chain = max(max_chain, chain);
    if (dense && !failed) {  // sorted form back; the region ends all zero
      if (l >= 0) {
        if (!huge_write_back(d, gb, summ, bits, lo, ww0, sh_minw, ((l - lo) >> 5) - ww0, &a_off, &a_cap, &a_len,
                             red, &sh_flag, &sh_off, &sh_cap))
          failed = true;
      } else {
        a_len = 0;  // every row cancelled: the window and the tail are zero
      }
    }
    if (tid == 0) {
      if (failed) {
        d.len[c] = 0;
        d.low[c] = -1;
        d.state[c] = ST_ZERO;
      } else {
        d.off[c] = a_off;
        d.cap[c] = a_cap;
        d.len[c] = a_len;
        d.low[c] = a_len ? l : -1;
        d.state[c] = a_len ? ST_PENDING : ST_ZERO;
        if (a_len && l >= minrow) {
          const uint32_t p = d.piv[l];
          if (p == kEmpty || p > (uint32_t)c) {
            const uint32_t pos = atomicAdd(&d.ctr[cw::kW + nx], 1u);
            d.work[nx][pos] = c;
          }
        }
      }
    }
    __syncthreads();
  }
  if (tid == 0 && n_add) {
    atomicAdd(&d.stats[0], n_add);
    atomicAdd(&d.stats[2], n_tiled);
    atomicAdd(&d.stats[5], n_add - n_tiled);
    atomicAdd(&d.stats[6], n_bytes);
    atomicMax(&d.ctr[cw::kChain], max_chain);
    atomicMax(&d.ctr[cw::kChain + 1], max_len);
  }
}

// Twist clearing at the end of a dimension (specseq.h:1039-1046).  A low row without a column
// (cells >= ncols of a prebuilt matrix) has nothing to clear.
__global__ void k_twist(Dev d, const int32_t* cols, int32_t ncols) {
  GRID_STRIDE(i, ncols) {
    const int32_t c = cols[i];
    if (d.state[c] == ST_PIVOT) {
      const int32_t l = d.low[c];
      if (l < d.ncols && d.state[l] == ST_PENDING) d.state[l] = ST_ZERO;
    }
  }
}


// ------------------------------------------------------------------------------------
// dualize (cohomology): the anti-transpose.  Dual cell n-1-c has boundary
// {n-1-x : c is a facet of x} and dimension maxdim - dim(c) (specseq.h:303-363,
// kernel.h:521-580); its pairs map back as (b', d') -> (n-1-d', n-1-b').
// ------------------------------------------------------------------------------------
__global__ void k_dual_count(int64_t n, const int64_t* colptr, const int32_t* rows, const uint8_t* removed,
                             int32_t* cnt) {
  GRID_STRIDE(c, n) {
    if (removed[c]) continue;
    for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) atomicAdd(&cnt[n - 1 - rows[k]], 1);
  }
}

__global__ void k_dual_fill(int64_t n, const int64_t* colptr, const int32_t* rows, const uint8_t* removed,
                            const int64_t* dcolptr, int32_t* fill, int32_t* drows) {
  GRID_STRIDE(c, n) {
    if (removed[c]) continue;
    for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) {
      const int64_t dc = n - 1 - rows[k];
      const int32_t pos = atomicAdd(&fill[dc], 1);
      drows[dcolptr[dc] + pos] = (int32_t)(n - 1 - c);
    }
  }
}

__global__ void k_dual_dims(int64_t n, const int32_t* dims, int32_t maxd, const uint8_t* removed, int32_t* ddims,
                            uint8_t* dremoved) {
  GRID_STRIDE(c, n) {
    ddims[n - 1 - c] = maxd - dims[c];
    dremoved[n - 1 - c] = removed[c];
  }
}

// ---- anti-transpose by a stable radix sort ---------------------------------------------
// Entry k = (row r, column c) becomes (key n-1-r, value n-1-c), written at position nnz-1-k:
// a stable sort by key then lists each dual column's rows in ascending order (larger c
// first).  Columns marked removed get key n and sort past every dual column.
__global__ void k_dual_keys(int64_t n, int64_t nnz, const int64_t* colptr, const int32_t* rows,
                            const uint8_t* removed, uint32_t* keys, int32_t* vals) {
  GRID_STRIDE(c, n) {
    const bool gone = removed && removed[c];
    const int32_t v = (int32_t)(n - 1 - c);
    for (int64_t k = colptr[c], ke = colptr[c + 1]; k < ke; k++) {
      const int64_t pos = nnz - 1 - k;
      keys[pos] = gone ? (uint32_t)n : (uint32_t)(n - 1 - rows[k]);
      vals[pos] = v;
    }
  }
}

// Runs of equal keys in the sorted keys: first[k] = start of key k's run (only where key k
// occurs); then len[k] = run length (len[] zeroed beforehand, so absent keys stay 0).  Each
// thread takes four consecutive keys in one 16-byte load (the key buffer is 16-byte aligned);
// 0xffffffff stands for "no neighbour" (keys are at most n < 2^31).
__device__ __forceinline__ void load_keys4(const uint32_t* keys, int64_t nnz, int64_t i0, uint32_t (&k)[4]) {
  if (i0 + 3 < nnz) {
    const uint4 v = *(const uint4*)(keys + i0);
    k[0] = v.x, k[1] = v.y, k[2] = v.z, k[3] = v.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) k[j] = i0 + j < nnz ? keys[i0 + j] : 0xffffffffu;
  }
}

__global__ void k_run_first(int64_t n, int64_t nnz, const uint32_t* keys, int64_t* first) {
  GRID_STRIDE(q, (nnz + 3) / 4) {
    const int64_t i0 = 4 * q;
    uint32_t k[4];
    load_keys4(keys, nnz, i0, k);
    uint32_t prev = i0 > 0 ? keys[i0 - 1] : 0xffffffffu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      if (k[j] < (uint64_t)n && k[j] != prev) first[k[j]] = i0 + j;
      prev = k[j];
    }
  }
}

__global__ void k_run_len(int64_t n, int64_t nnz, const uint32_t* keys, const int64_t* first, int32_t* len) {
  GRID_STRIDE(q, (nnz + 3) / 4) {
    const int64_t i0 = 4 * q;
    uint32_t k[4];
    load_keys4(keys, nnz, i0, k);
    const uint32_t next = i0 + 4 < nnz ? keys[i0 + 4] : 0xffffffffu;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint32_t nx = j < 3 ? k[j + 1] : next;
      if (k[j] < (uint64_t)n && k[j] != nx) len[k[j]] = (int32_t)(i0 + j + 1 - first[k[j]]);
    }
  }
}

__device__ void heapsort_i32(int32_t* a, int64_t n);

// Sort every segment v[off[s] : off[s+1]) ascending.  Segments longer than a lane's share
// are sorted by a wavefront: bitonic in registers up to 64 entries, in LDS up to kSortLds,
// heapsort by one lane beyond (rare).
constexpr int kSortLds = 2048;
__device__ void heapsort_i32(int32_t* a, int64_t n) {
  auto sift = [&](int64_t i, int64_t m) {
    for (;;) {
      int64_t l = 2 * i + 1, big = i;
      if (l < m && a[l] > a[big]) big = l;
      if (l + 1 < m && a[l + 1] > a[big]) big = l + 1;
      if (big == i) return;
      const int32_t t = a[i];
      a[i] = a[big];
      a[big] = t;
      i = big;
    }
  };
  for (int64_t i = n / 2 - 1; i >= 0; i--) sift(i, n);
  for (int64_t m = n - 1; m > 0; m--) {
    const int32_t t = a[0];
    a[0] = a[m];
    a[m] = t;
    sift(0, m);
  }
}

// A wavefront takes 64 consecutive segments: each lane sorts its own segment in registers
// when it has at most kSortShort entries (fixed compare-exchange network); the longer
// segments of the group are then sorted one after another by the whole wavefront.
constexpr int kSortShort = 16;
__global__ void __launch_bounds__(256) k_sort_segments(int64_t nseg, const int64_t* off, int32_t* v) {
  __shared__ int32_t lds[4][kSortLds];
  const int w = threadIdx.x / kWave, lane = lane_id();
  int32_t* buf = lds[w];
  for (int64_t g0 = ((int64_t)blockIdx.x * 4 + w) * kWave; g0 < nseg; g0 += (int64_t)gridDim.x * 4 * kWave) {
    const int64_t s0 = g0 + lane;
    int64_t mylen = 0;
    if (s0 < nseg) {
      const int64_t b = off[s0];
      mylen = off[s0 + 1] - b;
      if (mylen > 1 && mylen <= kSortShort) {
        int32_t x[kSortShort];
        for (int i = 0; i < kSortShort; i++) x[i] = i < mylen ? v[b + i] : 0x7fffffff;
        for (int i = 1; i < kSortShort; i++)  // fixed trip counts keep x[] in registers
          for (int j = i; j > 0; j--) {
            const int32_t a = x[j - 1], c = x[j];
            x[j - 1] = min(a, c);
            x[j] = max(a, c);
          }
        for (int i = 0; i < kSortShort; i++)
          if (i < mylen) v[b + i] = x[i];
      }
    }
    uint64_t longs = __ballot(mylen > kSortShort);
    while (longs) {
      const int src = __ffsll((unsigned long long)longs) - 1;
      longs &= longs - 1;
      const int64_t s = g0 + src;
      const int64_t b = off[s];
      const int64_t len = off[s + 1] - b;
      if (len <= kWave) {
        int32_t x = lane < len ? v[b + lane] : 0x7fffffff;
        for (int k = 2; k <= kWave; k <<= 1)
          for (int j = k >> 1; j > 0; j >>= 1) {
            const int32_t y = __shfl_xor(x, j, kWave);
            const bool up = (lane & k) == 0, lower = (lane & j) == 0;
            x = (lower == up) ? min(x, y) : max(x, y);
          }
        if (lane < len) v[b + lane] = x;
      } else if (len <= kSortLds) {
        int P = kWave;
        while (P < len) P <<= 1;
        for (int i = lane; i < P; i += kWave) buf[i] = i < len ? v[b + i] : 0x7fffffff;
        wave_sync();
        for (int k = 2; k <= P; k <<= 1)
          for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < P; i += kWave) {
              const int l = i ^ j;
              if (l > i) {
                const int32_t a = buf[i], c = buf[l];
                const bool asc = (i & k) == 0;
                if ((a > c) == asc) {
                  buf[i] = c;
                  buf[l] = a;
                }
              }
            }
            wave_sync();
          }
        for (int i = lane; i < len; i += kWave) v[b + i] = buf[i];
        wave_sync();
      } else if (lane == 0) {
        heapsort_i32(v + b, len);
      }
    }
  }
}

// exclusive scan of int32 counts into int64 offsets out[0..n] (out[n] = total)
// Exclusive scan of int32 counts into int64 offsets (out[n] = total) in three passes: tile
// sums, a scan of the tile sums, then each tile scanned with its base.  A tile is kScanTile
// counts: 16 per thread, loaded coalesced and staged through LDS (one pad slot per 16
// entries keeps a thread's 16 consecutive entries off a single bank).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanThreads * kScanPer;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 1; k < kWave; k <<= 1) {
    const int64_t y = __shfl_up(v, k, kWave);
    if (lane >= k) v += y;
  }
  return v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_partial(const int32_t* in, int64_t n, int64_t* bsum) {
  __shared__ int64_t red[kScanThreads / kWave];
  const int64_t b0 = (int64_t)blockIdx.x * kScanTile;
  int64_t acc = 0;
#pragma unroll
  for (int j = 0; j < kScanPer; j++) {
    const int64_t i = b0 + j * kScanThreads + threadIdx.x;
    if (i < n) acc += in[i];
  }
  for (int k = kWave / 2; k > 0; k >>= 1) acc += __shfl_xor(acc, k, kWave);
  if (lane_id() == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
    for (int w = 0; w < kScanThreads / kWave; w++) t += red[w];
    bsum[blockIdx.x] = t;
  }
}

// In place: bsum[0..nb) becomes its exclusive scan, bsum[nb] the total (one workgroup).
__global__ void __launch_bounds__(1024) k_scan_bsums(int64_t nb, int64_t* bsum) {
  __shared__ int64_t wsum[1024 / kWave];
  const int64_t per = (nb + 1023) / 1024;
  const int64_t a = (int64_t)threadIdx.x * per;
  const int64_t e = a + per < nb ? a + per : nb;
  int64_t sum = 0;
  for (int64_t b = a; b < e; b++) sum += bsum[b];
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const int64_t incl = wave_incl_scan(sum);
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    const int64_t v = lane < 1024 / kWave ? wsum[lane] : 0;
    const int64_t vi = wave_incl_scan(v);
    if (lane < 1024 / kWave) wsum[lane] = vi - v;
    if (lane == 1024 / kWave - 1) bsum[nb] = vi;
  }
  __syncthreads();
  int64_t run = wsum[w] + incl - sum;
  for (int64_t b = a; b < e; b++) {
    const int64_t v = bsum[b];
    bsum[b] = run;
    run += v;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_final(const int32_t* in, int64_t n, const int64_t* bsum,
                                                             int64_t nb, int64_t* out) {
  __shared__ int64_t s[kScanTile + kScanTile / kScanPer];
  __shared__ int64_t wsum[kScanThreads / kWave];
  const int64_t b0 = (int64_t)blockIdx.x * kScanTile;
#pragma unroll
  for (int j = 0; j < kScanPer; j++) {
    const int idx = j * kScanThreads + threadIdx.x;
    const int64_t i = b0 + idx;
    s[idx + idx / kScanPer] = i < n ? in[i] : 0;
  }
  __syncthreads();
  const int base = threadIdx.x * (kScanPer + 1);  // padded position of the thread's 16 entries
  int64_t x[kScanPer];
  int64_t tot = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    x[k] = s[base + k];
    tot += x[k];
  }
  const int lane = lane_id(), w = threadIdx.x / kWave;
  const int64_t incl = wave_incl_scan(tot);
  if (lane == kWave - 1) wsum[w] = incl;
  __syncthreads();
  int64_t run = bsum[blockIdx.x] + incl - tot;
  for (int ww = 0; ww < w; ww++) run += wsum[ww];
#pragma unroll
  for (int k = 0; k < kScanPer; k++) {
    s[base + k] = run;
    run += x[k];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kScanPer; j++) {
    const int idx = j * kScanThreads + threadIdx.x;
    const int64_t i = b0 + idx;
    if (i < n) out[i] = s[idx + idx / kScanPer];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = bsum[nb];
}

// death_of[b] for the primal pairing: every dual pair (r, piv[r]) is the primal pair
// (n-1-piv[r], n-1-r); death_of is -1 beforehand.
__global__ void k_dual_pairs_fill(int64_t n, const uint32_t* piv, int32_t* death_of) {
  GRID_STRIDE(r, n) {
    const uint32_t c = piv[r];
    if (c != kEmpty) death_of[n - 1 - (int64_t)c] = (int32_t)(n - 1 - r);
  }
}

__global__ void k_dual_essential(const int64_t* ne_dev, const int64_t* in, int64_t n, int64_t* out) {
  const int64_t ne = *ne_dev;
  GRID_STRIDE(k, ne) out[ne - 1 - k] = n - 1 - in[k];
}

__global__ void k_rowstat_init(int32_t* rowstat) {
  if (threadIdx.x < kMaxDims) {
    rowstat[threadIdx.x] = -1;
    rowstat[kMaxDims + threadIdx.x] = 0x7fffffff;
  }
}


// Block-diagonal batch: complex i's cells are shifted by base[i] and its nonzeros by
// nzb[i]; the host only concatenates the arrays unshifted.
__device__ __forceinline__ int64_t find_block(const int64_t* off, int64_t count, int64_t x) {
  int64_t lo = 0, hi = count;  // largest i with off[i] <= x
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (off[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Validation of host-supplied blocks before anything indexes with them: per column, the
// offsets lie in its block's row array and its rows are strictly increasing and below the
// block's cell count.  Offsets and rows are still block-local (unshifted).
__global__ void k_check_blocks(int64_t n, int64_t count, const int64_t* base, const int64_t* nzb,
                               const int64_t* colptr, const int32_t* rows, uint32_t* bad) {
  GRID_STRIDE(g, n) {
    const int64_t i = count > 1 ? find_block(base, count, g) : 0;
    const int64_t c = g - base[i], n_i = base[i + 1] - base[i], nz_i = nzb[i + 1] - nzb[i];
    const int64_t s = colptr[g], e = c + 1 < n_i ? colptr[g + 1] : nz_i;
    bool ok = (c != 0 || s == 0) && s >= 0 && e >= s && e <= nz_i;
    if (ok) {
      const int32_t* R = rows + nzb[i];
      int64_t prev = -1;
      for (int64_t k = s; k < e; k++) {
        const int64_t r = R[k];
        if (r <= prev || r >= n_i) {
          ok = false;
          break;
        }
        prev = r;
      }
    }
    if (!ok) *bad = 1u;
  }
}

__global__ void k_batch_shift(int64_t n, int64_t nnz, int64_t count, const int64_t* base, const int64_t* nzb,
                              int64_t* colptr, int32_t* rows) {
  GRID_STRIDE(c, n) colptr[c] += nzb[find_block(base, count, c)];
  GRID_STRIDE(k, nnz) rows[k] += (int32_t)base[find_block(nzb, count, k)];
  if (blockIdx.x == 0 && threadIdx.x == 0) colptr[n] = nnz;
}

// One resident matrix into the batch buffers at cell offset `base` and nonzero offset `nzb`:
// offsets and rows shifted to global indices.  Entry n of the offsets equals the next block's
// entry 0, so consecutive launches on one stream write the same value there.
__global__ void k_gather_block(int64_t n, int64_t nnz, const int64_t* __restrict__ cp,
                               const int32_t* __restrict__ rows, const int32_t* __restrict__ dims, int64_t base,
                               int64_t nzb, int64_t* __restrict__ dcp, int32_t* __restrict__ drows,
                               int32_t* __restrict__ ddims) {
  GRID_STRIDE(c, n + 1) dcp[c] = cp[c] + nzb;
  GRID_STRIDE(c, n) ddims[c] = dims[c];
  GRID_STRIDE(k, nnz) drows[k] = rows[k] + (int32_t)base;
}

// ------------------------------------------------------------------------------------
// edge collapse support (model bit 1)
// ------------------------------------------------------------------------------------
__global__ void k_edge_ends(int64_t ne, const int32_t* edge_cols, const int64_t* colptr, const int32_t* rows,
                            int32_t* u, int32_t* v, uint32_t* bad) {
  GRID_STRIDE(i, ne) {
    const int32_t c = edge_cols[i];
    const int64_t o = colptr[c];
    if (colptr[c + 1] - o != 2) {
      atomicOr(bad, 1u);
      u[i] = v[i] = 0;
    } else {
      u[i] = rows[o];
      v[i] = rows[o + 1];
    }
  }
}

__global__ void k_mark_removed(int64_t ne, const int32_t* edge_cols, const uint8_t* kept, uint8_t* removed) {
  GRID_STRIDE(i, ne) removed[edge_cols[i]] = kept[i] ? 0 : 1;
}

// Removal propagates to cofaces, one dimension at a time (kernel.h:190-198).
__global__ void k_propagate(int64_t n, int32_t dim, const int32_t* dims, const int64_t* colptr,
                            const int32_t* rows, uint8_t* removed) {
  GRID_STRIDE(c, n) {
    if (dims[c] != dim) continue;
    for (int64_t k = colptr[c]; k < colptr[c + 1]; k++)
      if (removed[rows[k]]) {
        removed[c] = 1;
        break;
      }
  }
}

// ------------------------------------------------------------------------------------
// high-dimension guided clearing (model bit 0)
// ------------------------------------------------------------------------------------
constexpr int kMaxVerts = 8;  // vertices per cell handled by guided clearing (dim <= 6)

// Vertex set of each cell of dimension dim from its first two facets (kernel.h:121-188),
// stored sorted in vs[c*kMaxVerts ...].  Vertex ids are vertex ranks.
__global__ void k_vertex_sets(int64_t n, int32_t dim, const int32_t* dims, const int64_t* colptr,
                              const int32_t* rows, const int32_t* vrank, int32_t* vs, uint32_t* bad) {
  GRID_STRIDE(c, n) {
    if (dims[c] != dim) continue;
    int32_t* out = vs + c * kMaxVerts;
    if (dim == 0) {
      out[0] = vrank[c];
      continue;
    }
    const int64_t o = colptr[c];
    if (colptr[c + 1] - o != dim + 1 || dims[rows[o]] != dim - 1 || dims[rows[o + 1]] != dim - 1) {
      atomicOr(bad, 1u);
      for (int k = 0; k <= dim; k++) out[k] = -1;
      continue;
    }
    const int32_t* f0 = vs + (int64_t)rows[o] * kMaxVerts;
    const int32_t* f1 = vs + (int64_t)rows[o + 1] * kMaxVerts;
    // merge two sorted dim-vertex sets (they share dim-1 vertices)
    int i = 0, j = 0, k = 0;
    while ((i < dim || j < dim) && k <= dim) {
      int32_t x = i < dim ? f0[i] : 0x7fffffff;
      int32_t y = j < dim ? f1[j] : 0x7fffffff;
      if (x == y) {
        out[k++] = x;
        i++;
        j++;
      } else if (x < y) {
        out[k++] = x;
        i++;
      } else {
        out[k++] = y;
        j++;
      }
    }
    if (k != dim + 1 || i != dim || j != dim) {
      atomicOr(bad, 1u);
      for (int t = 0; t <= dim; t++) out[t] = -1;
    }
  }
}

struct Key128 {
  unsigned long long lo, hi;
};

__device__ __forceinline__ Key128 pack_key(const int32_t* v, int nv, int bits) {
  Key128 k{0ull, 0ull};
  for (int t = 0; t < nv; t++) {
    const int sh = t * bits;
    const unsigned long long x = (unsigned long long)(uint32_t)v[t];
    if (sh < 64) {
      k.lo |= x << sh;
      if (sh + bits > 64) k.hi |= x >> (64 - sh);
    } else {
      k.hi |= x << (sh - 64);
    }
  }
  return k;
}

__device__ __forceinline__ uint64_t hash_key(Key128 k) {
  uint64_t h = k.lo * 0x9E3779B97F4A7C15ull ^ (k.hi + 0x632BE59BD9B4E019ull) * 0xC2B2AE3D27D4EB4Full;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  return h;
}

// Hash of the max-dim cells that are still in the complex: key -> column.
__global__ void k_hash_insert(int64_t n, int32_t dim, const int32_t* dims, const uint8_t* removed,
                              const int32_t* vs, int bits, Key128* keys, int32_t* vals, uint64_t mask) {
  GRID_STRIDE(c, n) {
    if (dims[c] != dim || removed[c]) continue;
    const int32_t* v = vs + c * kMaxVerts;
    if (v[0] < 0) continue;
    const Key128 key = pack_key(v, dim + 1, bits);
    uint64_t h = hash_key(key) & mask;
    for (;;) {
      if (atomicCAS(&vals[h], -1, (int32_t)c) == -1) {
        keys[h] = key;
        break;
      }
      h = (h + 1) & mask;
    }
  }
}

__device__ __forceinline__ int32_t hash_find(const Key128* keys, const int32_t* vals, uint64_t mask, Key128 key) {
  uint64_t h = hash_key(key) & mask;
  for (;;) {
    const int32_t v = vals[h];
    if (v < 0) return -1;
    if (keys[h].lo == key.lo && keys[h].hi == key.hi) return v;
    h = (h + 1) & mask;
  }
}

// Max-dim cell s is positive if some vertex x adjacent to all of its vertices gives a
// (dim+1)-simplex whose other facets are all cells with index < s (kernel.h:203-292;
// Lemma: the youngest facet of a boundary cycle is positive).  Neighbours of the vertex
// with the fewest neighbours are walked from the highest id down, at most max_iter of them.
__global__ void k_hd_clear(int64_t n, int32_t dim, const int32_t* dims, const uint8_t* removed,
                           const int32_t* vs, int bits, const int64_t* adj_off, const int32_t* adj,
                           const Key128* keys, const int32_t* vals, uint64_t mask, int32_t max_iter,
                           uint8_t* hdclear, uint32_t* ncleared) {
  GRID_STRIDE(c, n) {
    if (dims[c] != dim || removed[c]) continue;
    const int32_t* v = vs + c * kMaxVerts;
    if (v[0] < 0) continue;
    const int nv = dim + 1;
    int best = 0;
    for (int t = 1; t < nv; t++)
      if (adj_off[v[t] + 1] - adj_off[v[t]] < adj_off[v[best] + 1] - adj_off[v[best]]) best = t;
    const int64_t b0 = adj_off[v[best]];
    int64_t p = adj_off[v[best] + 1] - 1;
    bool cleared = false;
    for (int32_t it = 0; it < max_iter && p >= b0 && !cleared; it++, p--) {
      const int32_t x = adj[p];
      bool ok = true;
      for (int t = 0; t < nv && ok; t++) {
        if (t == best) continue;
        if (v[t] == x) { ok = false; break; }
        // binary search x in adj list of v[t]
        int64_t lo = adj_off[v[t]], hi = adj_off[v[t] + 1];
        while (lo < hi) {
          const int64_t mid = (lo + hi) >> 1;
          if (adj[mid] < x) lo = mid + 1;
          else hi = mid;
        }
        ok = lo < adj_off[v[t] + 1] && adj[lo] == x;
      }
      if (!ok) continue;
      // tau = v + {x}; every facet other than s itself must exist with index < c
      int32_t tau[kMaxVerts + 1];
      int k = 0;
      bool put = false;
      for (int t = 0; t < nv; t++) {
        if (!put && x < v[t]) {
          tau[k++] = x;
          put = true;
        }
        tau[k++] = v[t];
      }
      if (!put) tau[k++] = x;
      bool all = true;
      for (int drop = 0; drop <= nv && all; drop++) {
        if (tau[drop] == x) continue;  // that facet is s itself
        int32_t f[kMaxVerts];
        int q = 0;
        for (int t = 0; t <= nv; t++)
          if (t != drop) f[q++] = tau[t];
        const int32_t col = hash_find(keys, vals, mask, pack_key(f, nv, bits));
        all = col >= 0 && col < c;
      }
      if (all) cleared = true;
    }
    if (cleared) {
      hdclear[c] = 1;
      atomicAdd(ncleared, 1u);
    }
  }
}

// adjacency (CSR over vertex ranks) of the edges still in the complex
__global__ void k_adj_count(int64_t ne, const int32_t* eu, const int32_t* ev, const int32_t* edge_cols,
                            const uint8_t* removed, int64_t* deg) {
  GRID_STRIDE(i, ne) {
    if (removed[edge_cols[i]]) continue;
    atomicAdd((unsigned long long*)&deg[eu[i]], 1ull);
    atomicAdd((unsigned long long*)&deg[ev[i]], 1ull);
  }
}
__global__ void k_adj_fill(int64_t ne, const int32_t* eu, const int32_t* ev, const int32_t* edge_cols,
                           const uint8_t* removed, const int64_t* off, unsigned long long* fill, int32_t* adj) {
  GRID_STRIDE(i, ne) {
    if (removed[edge_cols[i]]) continue;
    const int32_t a = eu[i], b = ev[i];
    adj[off[a] + (int64_t)atomicAdd(&fill[a], 1ull)] = b;
    adj[off[b] + (int64_t)atomicAdd(&fill[b], 1ull)] = a;
  }
}
__global__ void k_adj_sort(int64_t nv, const int64_t* off, int32_t* adj) {
  GRID_STRIDE(a, nv) {
    const int64_t b = off[a], e = off[a + 1];
    for (int64_t i = b + 1; i < e; i++) {
      const int32_t x = adj[i];
      int64_t j = i - 1;
      while (j >= b && adj[j] > x) {
        adj[j + 1] = adj[j];
        j--;
      }
      adj[j + 1] = x;
    }
  }
}
__global__ void k_vrank_ends(int64_t ne, const int32_t* vrank, int32_t* eu, int32_t* ev, uint32_t* bad) {
  GRID_STRIDE(i, ne) {
    const int32_t a = vrank[eu[i]], b = vrank[ev[i]];
    if (a < 0 || b < 0) atomicOr(bad, 1u);  // an edge endpoint that is not a vertex
    eu[i] = a < 0 ? 0 : a;
    ev[i] = b < 0 ? 0 : b;
  }
}

// ------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------
#define DCHK(x)                \
  do {                         \
    const int rc_ = (x);       \
    if (rc_) return rc_;       \
  } while (0)
#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      set_error(std::string(#x) + ": " + hipGetErrorString(e_));                   \
      return SSQ_EDEVICE;                                                          \
    }                                                                              \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// The additions' workgroup shape in the parameters (0 = default) is one the engine supports.
static bool shape_ok(const ssq_params& p) {
  return (p.add_waves == 0 || p.add_waves == 2 || p.add_waves == 4 || p.add_waves == 8) && p.long_grid >= 0 &&
         p.long_grid <= 4096 && (p.short_max == 0 || (p.short_max >= 64 && p.short_max <= kTile)) &&
         p.rounds_per_sync >= 0 && p.rounds_per_sync <= 8;
}

class Engine : public PrebuiltTarget {
 public:
  explicit Engine(int device) : device_(device) {}
  ~Engine() override { release(); }

  int init() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
      err_ = "no HIP device";
      return SSQ_EDEVICE;
    }
    if (device_ < 0 || device_ >= ndev) {
      err_ = "bad device index";
      return SSQ_EINVAL;
    }
    HIPCHK(hipSetDevice(device_));
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc(&host_ctr_, cw::kWords * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHK(hipHostMalloc(&host_tot_, kTotSlots * sizeof(int64_t), hipHostMallocDefault));
    HIPCHK(hipMalloc(&tot_, kTotSlots * sizeof(int64_t)));
    HIPCHK(hipMalloc(&ctr_, cw::kWords * sizeof(uint32_t)));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device_));
    cus_ = prop.multiProcessorCount;
    // add waves resident per CU: the LDS staging buffers, not registers, set the limit
    int nb = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_add<2>, kWave * 2, 0));
    add_waves_per_cu_[0] = std::max(1, nb) * 2;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_add<4>, kWave * 4, 0));
    add_waves_per_cu_[1] = std::max(1, nb) * 4;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_add<8>, kWave * 8, 0));
    add_waves_per_cu_[2] = std::max(1, nb) * 8;
    const char* kt = std::getenv("SSQ_KERNEL_TIMING");
    timing_ = kt && kt[0] == '1';
    const char* tr = std::getenv("SSQ_TRACE_ROUNDS");  // per-round diagnostics on stderr
    trace_ = tr && tr[0] == '1';
    if (const char* e = std::getenv("SSQ_ROUNDS_PER_SYNC")) rounds_per_sync_ = std::max(1, std::min(8, std::atoi(e)));
    if (const char* e = std::getenv("SSQ_TEST_NO_SORT_DUAL")) no_sort_dual_ = e[0] == '1';
    // test hooks: smaller limits push short columns through the long-column paths
    if (const char* e = std::getenv("SSQ_TEST_SHORT_MAX")) short_max_ = std::max(2, std::min(kTile, std::atoi(e)));
    if (const char* e = std::getenv("SSQ_TEST_WINDOW_BITS"))
      long_window_ = std::max(64, std::min(kLongBits, std::atoi(e)));
    if (const char* e = std::getenv("SSQ_TEST_GLOBAL_BITS"))  // small regions push columns to tiled merges
      gspan_ = std::max(64, std::min(kHugeBits, std::atoi(e) & ~31));
    if (timing_)
      for (auto& e : ev_) HIPCHK(hipEventCreate(&e));
    return SSQ_OK;
  }

  const std::string& error() const { return err_; }
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }
  // byte fill and device-to-device copy on the engine's stream, by kernels
  int fill(void* dst, int value, size_t bytes) { return dfill(dst, value, bytes); }
  int copy(void* dst, const void* src, size_t bytes) { return dcopy(dst, src, bytes); }

  int reduce_device(int64_t n, int64_t nnz, const int64_t* d_colptr, const int32_t* d_rows,
                    const int32_t* d_dims, const ssq_params& p, int64_t* d_pairs, int64_t* np_out,
                    ssq_stats* st, int64_t* d_essential = nullptr, int64_t* ne_out = nullptr);
  // The same in the reference's two phases (GPU_Init, then GPU_Compute and the pairs): the engine
  // holds one initialized matrix between them; any other reduction on the engine discards it.
  int init_phase(int64_t n, int64_t nnz, const int64_t* d_colptr, const int32_t* d_rows, const int32_t* d_dims,
                 const ssq_params& p, ssq_stats* st);
  int compute_phase(int64_t* d_pairs, int64_t* np_out, ssq_stats* st, int64_t* d_essential = nullptr,
                    int64_t* ne_out = nullptr);
  // Prebuilt matrices: a builder writes the rows and the init tables straight into the engine's
  // buffers, then runs the init passes and the page loop on them.  Cells [0, n) are rows;
  // cells [0, ncols) also have columns (colptr[ncols + 1] into the pool, dims[ncols]; both stay
  // the caller's).  prebuilt_begin sizes the buffers and returns:
  //   pool[0, nnz): the rows to write; pool[nnz, nnz + arena): free (the store's arena, usable as
  //     scratch until prebuilt_reduce);
  //   leftone[n]: every row's leftmost column (kEmpty where none), to write;
  //   negflag[n]: zeroed; set to 1 for every column that is leftmost in some row.
  struct PrebuiltBufs {
    int32_t* pool;
    uint32_t* leftone;
    uint8_t* negflag;
  };
  int prebuilt_begin(int64_t n, int64_t ncols, int64_t nnz, int64_t arena, const int64_t* colptr, const int32_t* dims,
                     int32_t maxdim, PrebuiltBufs* out);
  // Reduces the prebuilt matrix.  *oom: the arena ran out (rebuild with a larger one).  After
  // it, low/state per column and the row -> column pivot table describe the pairs.
  int prebuilt_reduce(const ssq_params& p, ssq_stats* st, bool* oom);
  // PrebuiltTarget
  int pre_begin(int64_t n, int64_t ncols, int64_t nnz, int64_t arena, const int64_t* colptr, const int32_t* dims,
                int32_t maxdim, int32_t** pool, uint32_t** leftone, uint8_t** negflag) override {
    PrebuiltBufs b{};
    const int rc = prebuilt_begin(n, ncols, nnz, arena, colptr, dims, maxdim, &b);
    *pool = b.pool;
    *leftone = b.leftone;
    *negflag = b.negflag;
    return rc;
  }
  int pre_reduce(const ssq_params& p, ssq_stats* st, bool* oom) override { return prebuilt_reduce(p, st, oom); }
  const int32_t* pre_low() const override { return pre_.low; }
  const uint32_t* pre_piv() const override { return pre_.piv; }
  hipStream_t pre_stream() const override { return stream_; }
  int pre_device() const override { return device_; }
  const std::string& pre_error() const override { return err_; }

 private:
  void fill_stats(ssq_stats* S, int64_t nnz, int64_t arena_cap);
  void set_error(const std::string& s) { err_ = s; }
  template <typename T>
  int ensure(DevBuf& b, size_t count, T** out) {
    const size_t need = std::max<size_t>(count, 1) * sizeof(T);
    if (b.bytes < need) {
      if (b.p) (void)hipFree(b.p);
      b.p = nullptr;
      b.bytes = 0;
      const size_t alloc = need + need / 2;  // geometric growth: a free synchronizes the device
      if (hipMalloc(&b.p, alloc) != hipSuccess) {
        (void)hipGetLastError();
        err_ = "device allocation of " + std::to_string(alloc) + " bytes failed";
        return SSQ_ENOMEM;
      }
      b.bytes = alloc;
    }
    *out = (T*)b.p;
    return SSQ_OK;
  }
  void release() {
    for (DevBuf* b : {&b_pool_, &b_off_, &b_len_, &b_cap_, &b_low_, &b_state_, &b_removed_, &b_flag_, &b_piv_,
                      &b_left_, &b_work_, &b_misc_, &b_U_, &b_keys_, &b_cnt_, &b_e_, &b_vs_, &b_adj_,
                      &b_hash_, &b_hd_, &b_dcol_, &b_drow_, &b_ddim_, &b_drem_, &b_scan_, &b_pmeta_, &b_prem_,
                      &b_scrc_, &b_kbase_, &b_skey_, &b_stmp_, &b_gbits_})
      if (b->p) (void)hipFree(b->p), b->p = nullptr, b->bytes = 0;
    if (host_ctr_) (void)hipHostFree(host_ctr_), host_ctr_ = nullptr;
    if (host_tot_) (void)hipHostFree(host_tot_), host_tot_ = nullptr;
    if (tot_) (void)hipFree(tot_), tot_ = nullptr;
    if (ctr_) (void)hipFree(ctr_), ctr_ = nullptr;
    for (auto& e : ev_)
      if (e) (void)hipEventDestroy(e), e = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_), stream_ = nullptr;
  }
  int sync_read(const uint32_t* dsrc, int count) {
    HIPCHK(hipMemcpyAsync(host_ctr_, dsrc, sizeof(uint32_t) * count, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return SSQ_OK;
  }
  int grid_for(int64_t n, int block = 256) const {
    int64_t g = (n + block - 1) / block;
    const int64_t cap = (int64_t)cus_ * 16;
    g = std::min(g, cap);
    if (g >= kXcds) g = (g + kXcds - 1) / kXcds * kXcds;  // XCD-contiguous element loops
    return (int)std::max<int64_t>(1, g);
  }
  // Stable partition by a small key; totals go to device slot tot_[slot .. slot+nkeys).
  // With totals_host, they are also copied back (one sync); without, nothing syncs.
  template <typename KeyF, typename EmitF>
  int compact(int64_t n, int nkeys, KeyF key, EmitF emit, int64_t* totals_host, int slot = kTotScratch);
  // read counters and all total slots in one round trip
  // dst[0, bytes) = the byte value, on the engine's stream (k_fill_bytes)
  int dfill(void* dst, int value, size_t bytes) {
    if (bytes == 0) return SSQ_OK;
    const uintptr_t a = (uintptr_t)dst;
    const int64_t head = (int64_t)std::min<size_t>(bytes, (16 - (a & 15)) & 15);
    const int64_t words = (int64_t)((bytes - (size_t)head) / 16);
    const uint32_t b = (uint32_t)(value & 0xff);
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((std::max<int64_t>(words, 16) + 255) / 256, (int64_t)cus_ * 16));
    hipLaunchKernelGGL(k_fill_bytes, dim3((unsigned)g), dim3(256), 0, stream_, (uint8_t*)dst, b * 0x01010101u, head,
                       words, (int64_t)bytes);
    HIPCHK(hipGetLastError());
    return SSQ_OK;
  }
  // device-to-device copy of bytes on the engine's stream (k_copy_bytes); the ranges must not overlap
  int dcopy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SSQ_OK;
    const uintptr_t a = (uintptr_t)dst, b = (uintptr_t)src;
    int64_t head = 0, words = 0;
    if ((a & 15) == (b & 15)) {
      head = (int64_t)std::min<size_t>(bytes, (16 - (a & 15)) & 15);
      words = (int64_t)((bytes - (size_t)head) / 16);
    }
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>((std::max<int64_t>(words, 16) + 255) / 256, (int64_t)cus_ * 16));
    hipLaunchKernelGGL(k_copy_bytes, dim3((unsigned)g), dim3(256), 0, stream_, (uint8_t*)dst, (const uint8_t*)src, head,
                       words, (int64_t)bytes);
    HIPCHK(hipGetLastError());
    return SSQ_OK;
  }
  int sync_all() {
    HIPCHK(hipMemcpyAsync(host_ctr_, ctr_dev_, sizeof(uint32_t) * cw::kWords, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipMemcpyAsync(host_tot_, tot_, sizeof(int64_t) * kTotSlots, hipMemcpyDeviceToHost, stream_));
    HIPCHK(hipStreamSynchronize(stream_));
    return SSQ_OK;
  }
  int collapse(Dev& d, int64_t n, const int32_t* d_dims, const int64_t* d_colptr, const int32_t* d_rows,
               ssq_stats* st);
  int guided_clear(Dev& d, int64_t n, int32_t maxdim, const int32_t* d_dims, const int64_t* d_colptr,
                   const int32_t* d_rows, const ssq_params& p, uint8_t** hd_out, ssq_stats* st);
  int scan_counts(const int32_t* cnt, int64_t n, int64_t* out);
  struct LoopPlan {
    int32_t* U = nullptr;  // pending columns, grouped by dimension
    int64_t cnt[kMaxDims] = {}, ubase[kMaxDims] = {}, upend = 0;
    int32_t maxrow[kMaxDims] = {}, minrow[kMaxDims] = {};
  };
  int init_lists(Dev& d, int32_t maxdim, const uint8_t* hd, bool neg_known, const ssq_params& p, LoopPlan* L,
                 ssq_stats* S);
  // A reduction between its two phases: the matrix as reduced (the anti-transpose when dualize
  // is set), the column state, the pending lists and the counters so far.
  struct Phase {
    bool ready = false;
    Dev d{};
    LoopPlan L;
    ssq_params p{};
    int32_t maxdim = 0;
    bool dual = false;
    int64_t n = 0, nnz_eff = 0, arena_cap = 0;
    const int32_t* rows = nullptr;  // the rows the column store starts from
    const int64_t* primal_colptr = nullptr;
    const int32_t* primal_rows = nullptr;
    uint8_t* primal_removed = nullptr;
    uint8_t* hd = nullptr;
    int32_t* U = nullptr;
    ssq_stats S{};
    std::chrono::steady_clock::time_point T0, T1;
  };
  Phase ph_;
  int prepare_store();
  int page_loop(Dev& d, const ssq_params& p, int32_t maxdim, const LoopPlan& L, ssq_stats* S, bool* oom_out);
  // The huge-column bit set regions, allocated and zeroed once (k_add_huge leaves them zero).
  int ensure_gbits(Dev* d) {
    const size_t words = (size_t)kHugeGroups * (size_t)(gspan_ >> 5);
    if (!gbits_ready_) {
      uint32_t* g;
      int rc = ensure(b_gbits_, words, &g);
      if (rc) return rc;
      DCHK(dfill(g, 0, words * sizeof(uint32_t)));
      gbits_ready_ = true;
    }
    d->gbits = (uint32_t*)b_gbits_.p;
    d->gspan = gspan_;
    return SSQ_OK;
  }
  DevBuf b_gbits_;
  bool gbits_ready_ = false;
  int32_t gspan_ = kHugeBits;
  Dev pre_{};                // the prebuilt matrix between prebuilt_begin and prebuilt_reduce
  int32_t pre_maxdim_ = 0;
  int build_dual(Dev& d, int64_t n, int64_t nnz, int32_t maxdim, const int64_t* colptr, const int32_t* rows,
                 const int32_t* dims, bool any_removed, int64_t** dcolptr, int32_t** drows, int32_t** ddims);
  int dual_by_sort(Dev& d, int64_t n, int64_t nnz, const int64_t* colptr, const int32_t* rows, bool any_removed,
                   int64_t* dcolptr, int32_t* drows);

  static constexpr int kTotPending = 0;   // [0..7] pending columns per dimension
  static constexpr int kTotPairs = 8;     // pairs
  static constexpr int kTotEss = 9;       // essential cells
  static constexpr int kTotDualPairs = 10;
  static constexpr int kTotScratch = 16;  // [16..23] counts only needed right away
  static constexpr int kTotSlots = 24;
  int64_t* tot_ = nullptr;       // device
  uint32_t* ctr_ = nullptr;      // device: the per-call counter block (cw::kWords words)
  int64_t* host_tot_ = nullptr;  // pinned
  uint32_t* ctr_dev_ = nullptr;

  int device_ = 0;
  int cus_ = 256;
  int add_waves_per_cu_[3] = {12, 12, 12};  // resident add wavefronts per CU with 2, 4, 8 per workgroup
  hipStream_t stream_ = nullptr;
  uint32_t* host_ctr_ = nullptr;
  std::string err_;
  DevBuf b_pool_, b_off_, b_len_, b_cap_, b_low_, b_state_, b_removed_, b_flag_, b_piv_, b_left_, b_work_,
      b_misc_, b_U_, b_keys_, b_cnt_, b_e_, b_vs_, b_adj_, b_hash_, b_hd_, b_dcol_, b_drow_, b_ddim_,
      b_drem_, b_scan_, b_pmeta_, b_prem_, b_scrc_, b_kbase_, b_skey_, b_stmp_;
  int64_t arena_hint_ = 0;
  bool no_sort_dual_ = false;  // test hook: always use the atomic anti-transpose
  bool timing_ = false;
  bool trace_ = false;
  int rounds_per_sync_ = 3;  // add rounds launched between two reads of the worklist counters
  int32_t short_max_ = kTile;
  int32_t long_window_ = kLongBits;
  hipEvent_t ev_[32] = {};  // 4 per round, up to 8 rounds per sync
};

template <typename KeyF, typename EmitF>
int Engine::compact(int64_t n, int nkeys, KeyF key, EmitF emit, int64_t* totals_host, int slot) {
  const int64_t nb = std::max<int64_t>(1, (n + kCompactItems - 1) / kCompactItems);
  uint32_t* counts;
  int64_t* bases;
  int rc;
  if ((rc = ensure(b_cnt_, (size_t)(nb * nkeys + 16), &counts)) || (rc = ensure(b_kbase_, (size_t)(nb * nkeys + 1), &bases)))
    return rc;
  int64_t* totals = tot_ + slot;
  if (n > 0) {
    hipLaunchKernelGGL(k_key_count<KeyF>, dim3((unsigned)nb), dim3(kCompactBlock), 0, stream_, n, key, counts,
                       nkeys);
  } else {
    DCHK(dfill(counts, 0, sizeof(uint32_t) * nkeys));
  }
  if ((rc = scan_counts((const int32_t*)counts, nb * nkeys, bases))) return rc;
  hipLaunchKernelGGL(k_key_totals, dim3(1), dim3(64), 0, stream_, nkeys, nb, (const int64_t*)bases, totals);
  if (n > 0)
    hipLaunchKernelGGL((k_key_scatter<KeyF, EmitF>), dim3((unsigned)nb), dim3(kCompactBlock), 0, stream_, n, key,
                       (const int64_t*)bases, nkeys, emit);
  HIPCHK(hipGetLastError());
  if (!totals_host) return SSQ_OK;  // the caller reads tot_[slot..] at its next sync
  HIPCHK(hipMemcpyAsync(host_tot_ + slot, totals, sizeof(int64_t) * nkeys, hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  std::memcpy(totals_host, host_tot_ + slot, sizeof(int64_t) * nkeys);
  return SSQ_OK;
}

struct KeyPendingDim {
  const uint8_t* state;
  const int32_t* dims;
  int32_t nkeys;  // keys are dimensions 1..nkeys-1
  __device__ int operator()(int64_t c) const {
    const int32_t d = dims[c];
    return (state[c] == ST_PENDING && d >= 1 && d < nkeys) ? d : -1;
  }
};
struct EmitIndex {
  int32_t* out;
  __device__ void operator()(int64_t i, int64_t pos) const { out[pos] = (int32_t)i; }
};
struct KeyDimEq {
  const int32_t* dims;
  int32_t dim;
  __device__ int operator()(int64_t c) const { return dims[c] == dim ? 0 : -1; }
};
struct KeyPivot {
  const uint32_t* piv;
  __device__ int operator()(int64_t r) const { return piv[r] != kEmpty ? 0 : -1; }
};
struct EmitPair {
  const uint32_t* piv;
  int64_t* out;
  __device__ void operator()(int64_t r, int64_t pos) const {
    out[2 * pos] = r;
    out[2 * pos + 1] = (int64_t)piv[r];
  }
};
struct KeyEssential {
  const uint32_t* piv;
  const uint8_t* removed;
  const uint8_t* state;
  const int32_t* low;
  __device__ int operator()(int64_t c) const {
    if (removed[c] || piv[c] != kEmpty) return -1;
    if (state[c] == ST_PIVOT && low[c] >= 0 && piv[low[c]] == (uint32_t)c) return -1;
    return 0;
  }
};
struct EmitI64 {
  int64_t* out;
  __device__ void operator()(int64_t i, int64_t pos) const { out[pos] = i; }
};
struct KeyDeathOf {
  const int32_t* death_of;
  __device__ int operator()(int64_t r) const { return death_of[r] >= 0 ? 0 : -1; }
};
struct EmitDeathOf {
  const int32_t* death_of;
  int64_t* out;
  __device__ void operator()(int64_t r, int64_t pos) const {
    out[2 * pos] = r;
    out[2 * pos + 1] = death_of[r];
  }
};
struct KeyVertex {
  const int32_t* dims;
  __device__ int operator()(int64_t c) const { return dims[c] == 0 ? 0 : -1; }
};
struct EmitRank {
  int32_t* vrank;
  __device__ void operator()(int64_t i, int64_t pos) const { vrank[i] = (int32_t)pos; }
};

int Engine::collapse(Dev& d, int64_t n, const int32_t* d_dims, const int64_t* d_colptr, const int32_t* d_rows,
                     ssq_stats* st) {
  int32_t* ecols;
  int rc = ensure(b_e_, (size_t)(3 * n + 16), &ecols);
  if (rc) return rc;
  int64_t ne = 0;
  rc = compact(n, 1, KeyDimEq{d_dims, 1}, EmitIndex{ecols}, &ne);
  if (rc) return rc;
  if (ne == 0) return SSQ_OK;
  int32_t* eu = ecols + n;
  int32_t* ev = ecols + 2 * n;
  DCHK(dfill(d.ctr + cw::kFlags, 0, sizeof(uint32_t)));
  hipLaunchKernelGGL(k_edge_ends, dim3(grid_for(ne)), dim3(256), 0, stream_, ne, ecols, d_colptr, d_rows, eu, ev,
                     d.ctr + cw::kFlags);
  std::vector<int32_t> hu((size_t)ne), hv((size_t)ne);
  std::vector<uint8_t> kept((size_t)ne);
  HIPCHK(hipMemcpyAsync(hu.data(), eu, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipMemcpyAsync(hv.data(), ev, sizeof(int32_t) * ne, hipMemcpyDeviceToHost, stream_));
  rc = sync_read(d.ctr + cw::kFlags, 1);
  if (rc) return rc;
  if (host_ctr_[0]) {
    err_ = "edge column without exactly two boundary vertices";
    return SSQ_EFORMAT;
  }
  auto t0 = std::chrono::steady_clock::now();
  int threads = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  collapse_edges(ne, hu.data(), hv.data(), kept.data(), threads);
  auto t1 = std::chrono::steady_clock::now();
  int64_t nk = 0;
  for (uint8_t k : kept) nk += k;
  if (st) {
    st->t_collapse_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    st->n_kept_edges += nk;
  }
  HIPCHK(hipMemcpyAsync(d.flag, kept.data(), (size_t)ne, hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(k_mark_removed, dim3(grid_for(ne)), dim3(256), 0, stream_, ne, ecols, d.flag, d.removed);
  int32_t maxdim = 0;
  for (int dd = 2; dd < kMaxDims; dd++) {
    hipLaunchKernelGGL(k_propagate, dim3(grid_for(n)), dim3(256), 0, stream_, n, dd, d_dims, d_colptr, d_rows,
                       d.removed);
    maxdim = dd;
  }
  (void)maxdim;
  HIPCHK(hipStreamSynchronize(stream_));  // kept[] is a host temporary
  HIPCHK(hipGetLastError());
  return SSQ_OK;
}

int Engine::guided_clear(Dev& d, int64_t n, int32_t maxdim, const int32_t* d_dims, const int64_t* d_colptr,
                         const int32_t* d_rows, const ssq_params& p, uint8_t** hd_out, ssq_stats* st) {
  *hd_out = nullptr;
  if (maxdim < 1 || maxdim + 1 > kMaxVerts) return SSQ_OK;
  uint8_t* hd;
  int rc = ensure(b_hd_, (size_t)n, &hd);
  if (rc) return rc;
  DCHK(dfill(hd, 0, (size_t)n));
  // vertex ranks
  int32_t* vs;
  rc = ensure(b_vs_, (size_t)(n * (kMaxVerts + 1) + 16), &vs);
  if (rc) return rc;
  int32_t* vrank = vs + n * kMaxVerts;
  DCHK(dfill(vrank, 0xff, sizeof(int32_t) * n));
  int64_t nv = 0;
  rc = compact(n, 1, KeyVertex{d_dims}, EmitRank{vrank}, &nv);
  if (rc) return rc;
  int bits = 1;
  while (bits < 31 && (1ll << bits) < std::max<int64_t>(nv, 2)) bits++;
  if ((int64_t)bits * (maxdim + 1) > 128) return SSQ_OK;  // keys would not fit 128 bits
  DCHK(dfill(d.ctr + cw::kFlags, 0, sizeof(uint32_t) * 2));
  for (int32_t dd = 0; dd <= maxdim; dd++)
    hipLaunchKernelGGL(k_vertex_sets, dim3(grid_for(n)), dim3(256), 0, stream_, n, dd, d_dims, d_colptr, d_rows,
                       vrank, vs, d.ctr + cw::kFlags);
  // edges (in column order) -> adjacency over vertex ranks, kept edges only
  int32_t* ecols;
  rc = ensure(b_e_, (size_t)(3 * n + 16), &ecols);
  if (rc) return rc;
  int64_t ne = 0;
  rc = compact(n, 1, KeyDimEq{d_dims, 1}, EmitIndex{ecols}, &ne);
  if (rc) return rc;
  int32_t* eu = ecols + n;
  int32_t* ev = ecols + 2 * n;
  hipLaunchKernelGGL(k_edge_ends, dim3(grid_for(std::max<int64_t>(ne, 1))), dim3(256), 0, stream_, ne, ecols,
                     d_colptr, d_rows, eu, ev, d.ctr + cw::kFlags + 1);
  hipLaunchKernelGGL(k_vrank_ends, dim3(grid_for(std::max<int64_t>(ne, 1))), dim3(256), 0, stream_, ne, vrank, eu,
                     ev, d.ctr + cw::kFlags + 1);
  rc = sync_read(d.ctr + cw::kFlags, 2);
  if (rc) return rc;
  if (host_ctr_[0] || host_ctr_[1]) return SSQ_OK;  // not a simplicial complex: no guided clearing
  int64_t* aoff;
  rc = ensure(b_adj_, (size_t)(2 * (nv + 1) + ne + 8), &aoff);
  if (rc) return rc;
  unsigned long long* fill = (unsigned long long*)(aoff + nv + 1);
  int32_t* adj = (int32_t*)(aoff + 2 * (nv + 1));
  DCHK(dfill(aoff, 0, sizeof(int64_t) * 2 * (nv + 1)));
  hipLaunchKernelGGL(k_adj_count, dim3(grid_for(std::max<int64_t>(ne, 1))), dim3(256), 0, stream_, ne, eu, ev, ecols,
                     d.removed, aoff + 1);
  // exclusive scan of degrees on the host (nv is the vertex count; small)
  std::vector<int64_t> deg((size_t)nv + 1);
  HIPCHK(hipMemcpyAsync(deg.data(), aoff, sizeof(int64_t) * (nv + 1), hipMemcpyDeviceToHost, stream_));
  HIPCHK(hipStreamSynchronize(stream_));
  for (int64_t a = 0; a < nv; a++) deg[a + 1] += deg[a];
  HIPCHK(hipMemcpyAsync(aoff, deg.data(), sizeof(int64_t) * (nv + 1), hipMemcpyHostToDevice, stream_));
  hipLaunchKernelGGL(k_adj_fill, dim3(grid_for(std::max<int64_t>(ne, 1))), dim3(256), 0, stream_, ne, eu, ev, ecols,
                     d.removed, aoff, fill, adj);
  hipLaunchKernelGGL(k_adj_sort, dim3(grid_for(nv)), dim3(256), 0, stream_, nv, aoff, adj);
  // hash of max-dim cells
  int64_t nmax = 0;
  {
    int64_t cnt[1];
    rc = compact(n, 1, KeyDimEq{d_dims, maxdim}, EmitIndex{ecols}, cnt);  // ecols reused as scratch
    if (rc) return rc;
    nmax = cnt[0];
  }
  uint64_t hsize = 1;
  while (hsize < (uint64_t)(2 * nmax + 2)) hsize <<= 1;
  Key128* keys;
  rc = ensure(b_hash_, (size_t)(hsize * (sizeof(Key128) + sizeof(int32_t))), (uint8_t**)&keys);
  if (rc) return rc;
  int32_t* vals = (int32_t*)(keys + hsize);
  DCHK(dfill(vals, 0xff, sizeof(int32_t) * hsize));
  hipLaunchKernelGGL(k_hash_insert, dim3(grid_for(n)), dim3(256), 0, stream_, n, maxdim, d_dims, d.removed, vs, bits,
                     keys, vals, (uint64_t)(hsize - 1));
  DCHK(dfill(d.ctr + cw::kHdCleared, 0, sizeof(uint32_t)));
  hipLaunchKernelGGL(k_hd_clear, dim3(grid_for(n)), dim3(256), 0, stream_, n, maxdim, d_dims, d.removed, vs, bits,
                     aoff, adj, keys, vals, (uint64_t)(hsize - 1), p.max_iteration, hd, d.ctr + cw::kHdCleared);
  HIPCHK(hipGetLastError());
  rc = sync_read(d.ctr + cw::kHdCleared, 1);
  if (rc) return rc;
  if (st) st->n_cleared_hd += host_ctr_[0];
  *hd_out = hd;
  return SSQ_OK;
}


int Engine::scan_counts(const int32_t* cnt, int64_t n, int64_t* out) {
  const int64_t nb = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
  int64_t* bsum;
  int rc = ensure(b_scan_, (size_t)(nb + 1), &bsum);
  if (rc) return rc;
  hipLaunchKernelGGL(k_scan_partial, dim3((unsigned)nb), dim3(kScanThreads), 0, stream_, cnt, n, bsum);
  hipLaunchKernelGGL(k_scan_bsums, dim3(1), dim3(1024), 0, stream_, nb, bsum);
  hipLaunchKernelGGL(k_scan_final, dim3((unsigned)nb), dim3(kScanThreads), 0, stream_, cnt, n, bsum, nb, out);
  HIPCHK(hipGetLastError());
  return SSQ_OK;
}

// Stable LSD radix sort of (uint32 key, int32 value) pairs over key bits [0, end_bit): rocPRIM's
// onesweep with R-bit digits per pass.
template <unsigned R>
static hipError_t sort_pairs_r(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                               int32_t* vout, int64_t count, int end_bit, hipStream_t s) {
  using Onesweep = rocprim::radix_sort_onesweep_config<rocprim::kernel_config<1024, 16>, rocprim::kernel_config<1024, 16>,
                                                       R, rocprim::block_radix_rank_algorithm::match>;
  using Cfg = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config, Onesweep>;
  return rocprim::radix_sort_pairs<Cfg>(tmp, bytes, kin, kout, vin, vout, (size_t)count, 0u, (unsigned)end_bit, s);
}
// 9-bit digits where they save a pass over 8-bit ones (keys of 17-18 or 25-27 bits: 3 passes
// instead of 4 on 124M pairs is 0.9 ms less); 8-bit passes are the faster ones otherwise.
static hipError_t sort_pairs(void* tmp, size_t& bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                             int32_t* vout, int64_t count, int end_bit, hipStream_t s) {
  if ((end_bit + 8) / 9 < (end_bit + 7) / 8)
    return sort_pairs_r<9>(tmp, bytes, kin, kout, vin, vout, count, end_bit, s);
  return sort_pairs_r<8>(tmp, bytes, kin, kout, vin, vout, count, end_bit, s);
}

// Anti-transpose by a stable radix sort of (n-1-r, n-1-c) pairs over the key's bits; the
// sorted values are the dual's rows, the runs of equal keys its column lengths.
int Engine::dual_by_sort(Dev& d, int64_t n, int64_t nnz, const int64_t* colptr, const int32_t* rows,
                         bool any_removed, int64_t* dcolptr, int32_t* drows) {
  const uint8_t* rem = any_removed ? (const uint8_t*)d.removed : nullptr;
  uint32_t* keys;
  int32_t* vals;
  int32_t* len;
  int rc;
  const int64_t m = std::max<int64_t>(nnz, 1);
  if ((rc = ensure(b_skey_, (size_t)(2 * m), &keys)) || (rc = ensure(b_scrc_, (size_t)m, &vals)) ||
      (rc = ensure(b_misc_, (size_t)n, &len)))
    return rc;
  uint32_t* keys_out = keys + m;
  int end_bit = 1;
  while (end_bit < 32 && ((int64_t)1 << end_bit) <= n) end_bit++;  // keys lie in [0, n]
  size_t tb = 0;
  if (sort_pairs(nullptr, tb, keys, keys_out, vals, drows, nnz, end_bit, stream_) != hipSuccess) {
    err_ = "radix sort sizing failed";
    return SSQ_EDEVICE;
  }
  uint8_t* tmp;
  if ((rc = ensure(b_stmp_, tb, &tmp))) return rc;
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_dual_keys, dim3(g), dim3(256), 0, stream_, n, nnz, colptr, rows, rem, keys, vals);
  HIPCHK(hipGetLastError());
  if (nnz > 0 && sort_pairs(tmp, tb, keys, keys_out, vals, drows, nnz, end_bit, stream_) != hipSuccess) {
    err_ = "radix sort failed";
    return SSQ_EDEVICE;
  }
  DCHK(dfill(len, 0, sizeof(int32_t) * n));
  const int ge = grid_for((nnz + 3) / 4);
  hipLaunchKernelGGL(k_run_first, dim3(ge), dim3(256), 0, stream_, n, nnz, keys_out, dcolptr);
  hipLaunchKernelGGL(k_run_len, dim3(ge), dim3(256), 0, stream_, n, nnz, keys_out, dcolptr, len);
  HIPCHK(hipGetLastError());
  return scan_counts(len, n, dcolptr);  // run starts are overwritten by the column pointers
}

int Engine::build_dual(Dev& d, int64_t n, int64_t nnz, int32_t maxdim, const int64_t* colptr, const int32_t* rows,
                       const int32_t* dims, bool any_removed, int64_t** dcolptr, int32_t** drows, int32_t** ddims) {
  int32_t* cnt;
  uint8_t* drem;
  int rc;
  if ((rc = ensure(b_dcol_, (size_t)(n + 1), dcolptr)) || (rc = ensure(b_drow_, (size_t)std::max<int64_t>(nnz, 1), drows)) ||
      (rc = ensure(b_ddim_, (size_t)n, ddims)) || (rc = ensure(b_drem_, (size_t)n, &drem)) ||
      (rc = ensure(b_misc_, (size_t)n, &cnt)))
    return rc;
  const int g = grid_for(n);
  bool sorted = false;
  if (!no_sort_dual_) {
    rc = dual_by_sort(d, n, nnz, colptr, rows, any_removed, *dcolptr, *drows);
    if (rc == SSQ_OK) sorted = true;
    else if (rc != SSQ_ENOMEM) return rc;  // no room for the sort's buffers: the atomic path
  }
  if (sorted) {
    hipLaunchKernelGGL(k_dual_dims, dim3(g), dim3(256), 0, stream_, n, dims, maxdim, d.removed, *ddims, drem);
    HIPCHK(hipGetLastError());
    DCHK(dcopy(d.removed, drem, (size_t)n));
    return SSQ_OK;
  }
  DCHK(dfill(cnt, 0, sizeof(int32_t) * n));
  hipLaunchKernelGGL(k_dual_count, dim3(g), dim3(256), 0, stream_, n, colptr, rows, d.removed, cnt);
  rc = scan_counts(cnt, n, *dcolptr);
  if (rc) return rc;
  DCHK(dfill(cnt, 0, sizeof(int32_t) * n));
  hipLaunchKernelGGL(k_dual_fill, dim3(g), dim3(256), 0, stream_, n, colptr, rows, d.removed, *dcolptr, cnt, *drows);
  const int64_t sort_groups = (n + 4 * kWave - 1) / (4 * kWave);
  hipLaunchKernelGGL(k_sort_segments, dim3((unsigned)std::max<int64_t>(1, std::min<int64_t>(sort_groups, (int64_t)cus_ * 32))),
                     dim3(256), 0, stream_, n, *dcolptr, *drows);
  hipLaunchKernelGGL(k_dual_dims, dim3(g), dim3(256), 0, stream_, n, dims, maxdim, d.removed, *ddims, drem);
  HIPCHK(hipGetLastError());
  DCHK(dcopy(d.removed, drem, (size_t)n));
  return SSQ_OK;
}

// Init passes on a matrix whose rows are in the column store and whose column state k_setup
// has written (leftone set; with neg_known also the negative flags): apparent pairs,
// clearing, compression, then the per-dimension pending lists, page row bounds and level
// slots, read back in one round trip.
int Engine::init_lists(Dev& d, int32_t maxdim, const uint8_t* hd, bool neg_known, const ssq_params& p, LoopPlan* L,
                       ssq_stats* S) {
  const int g = grid_for(d.ncols);
  hipLaunchKernelGGL(k_lookup, dim3(g), dim3(256), 0, stream_, d, hd, neg_known ? 1 : 0);
  int32_t* rowstat = (int32_t*)(d.ctr + cw::kRowStat);
  hipLaunchKernelGGL(k_rowstat_init, dim3(1), dim3(64), 0, stream_, rowstat);
  hipLaunchKernelGGL(k_clear_compress_stats, dim3(g), dim3(256), 0, stream_, d, rowstat, rowstat + kMaxDims);
  HIPCHK(hipGetLastError());
  // pending lists per dimension and page bounds, read back in one round trip
  const int pend_keys = std::min(kMaxDims, std::max(maxdim, 1) + 1);  // dims 1..maxdim hold pending columns
  DCHK(dfill(tot_ + kTotPending, 0, sizeof(int64_t) * kMaxDims));
  int rc = compact(d.ncols, pend_keys, KeyPendingDim{d.state, d.dims, pend_keys}, EmitIndex{L->U}, nullptr, kTotPending);
  if (rc) return rc;

  if ((rc = sync_all())) return rc;
  int64_t upend = 0;
  for (int k = 0; k < kMaxDims; k++) {
    L->cnt[k] = host_tot_[kTotPending + k];
    L->ubase[k] = upend;
    upend += L->cnt[k];
  }
  for (int k = 0; k < kMaxDims; k++) {
    L->maxrow[k] = (int32_t)host_ctr_[cw::kRowStat + k];
    L->minrow[k] = (int32_t)host_ctr_[cw::kRowStat + kMaxDims + k];
  }
  S->n_apparent = host_ctr_[cw::kApparent];
  S->n_pending = upend;
  L->upend = upend;
  // level slots (specseq.h:281-283), arena-backed
  d.lvl_cap[1] = p.level1;
  d.lvl_cap[2] = p.level2;
  d.lvl_cap[3] = (int32_t)std::min<int64_t>((int64_t)p.level2 * 8, 0x3fffffff);
  const int64_t U1 = upend;
  d.lvl_slots[1] = (uint32_t)std::min<int64_t>(U1 + 1, 0xffffffffLL);
  d.lvl_slots[2] = (uint32_t)std::min<int64_t>(U1 / std::max(1, p.level2 / p.level1) / 2 + 1, 0xffffffffLL);
  d.lvl_slots[3] = (uint32_t)std::min<int64_t>(U1 / std::max<int64_t>(1, (int64_t)d.lvl_cap[3] / p.level1) / 2 + 1,
                                               0xffffffffLL);
  return SSQ_OK;
}

// Claim/add rounds per dimension, high to low (specseq.h:598-704), each page until its
// worklist is empty, then twist clearing.  *oom_out: the column store ran out (the caller
// reruns with a larger arena).
int Engine::page_loop(Dev& d, const ssq_params& p, int32_t maxdim, const LoopPlan& L, ssq_stats* S, bool* oom_out) {
  // page loop, dimensions high to low (specseq.h:598-704)
  const int64_t upend = L.upend;
  int32_t* W0;
  int rc = ensure(b_work_, (size_t)(6 * std::max<int64_t>(upend, 1) + 64), &W0);
  if (rc) return rc;
  if ((rc = ensure_gbits(&d))) return rc;
  const int64_t wcap = std::max<int64_t>(upend, 1);
  for (int k = 0; k < 6; k++) d.work[k] = W0 + k * wcap;
  // workgroup shape: the parameters' choice, else the defaults
  const int wpg = p.add_waves > 0 ? p.add_waves : kAddWavesDefault;  // 2, 4 or 8 (checked)
  const int wsel = wpg == 2 ? 0 : wpg == 4 ? 1 : 2;
  const int add_waves = (int)std::min<int64_t>(p.queue_size, (int64_t)cus_ * add_waves_per_cu_[wsel]);
  const int add_grid = std::max(1, (add_waves + wpg - 1) / wpg);
  const int rounds_per_sync = p.rounds_per_sync > 0 ? p.rounds_per_sync : rounds_per_sync_;
  const int long_grid = p.long_grid > 0 ? p.long_grid : cus_ * 2;
  if (p.short_max > 0) d.short_max = std::min(p.short_max, (int32_t)kTile);
  bool oom = false;
  *oom_out = false;
  for (int32_t dim = maxdim; dim >= 1 && !oom; dim--) {
    const int64_t nl = L.cnt[dim];
    if (nl == 0 || L.maxrow[dim] < 0) continue;
    const int32_t* cols = L.U + L.ubase[dim];
    int32_t ss = p.ss_block_size;
    int64_t begin = 0, end = std::min<int64_t>(ss, nl);
    int64_t mrow = (int64_t)L.maxrow[dim] - ss, xrow = mrow + ss;
    const int64_t dminrow = L.minrow[dim];
    while ((begin < nl || xrow > dminrow) && !oom) {
      DCHK(dfill(d.ctr, 0, 8 * sizeof(uint32_t)));
      const int64_t wn = end - begin;
      if (wn > 0) {
        const int32_t fl = (int32_t)std::max<int64_t>(mrow, INT32_MIN);
        hipLaunchKernelGGL(k_scan_window, dim3(grid_for(wn)), dim3(256), 0, stream_, d, cols + begin, (int32_t)wn,
                           fl);
        const int cg = grid_for(wn);
        int r = 0;
        for (;;) {
          for (int t = 0; t < rounds_per_sync; t++, r++) {
            hipLaunchKernelGGL(k_claim, dim3(cg), dim3(256), 0, stream_, d, r);
            hipLaunchKernelGGL(k_pivmeta, dim3(cg), dim3(256), 0, stream_, d, r);
            if (timing_) HIPCHK(hipEventRecord(ev_[4 * t], stream_));
            if (wpg == 2)
              hipLaunchKernelGGL(k_add<2>, dim3(add_grid), dim3(kWave * 2), 0, stream_, d, r, fl);
            else if (wpg == 4)
              hipLaunchKernelGGL(k_add<4>, dim3(add_grid), dim3(kWave * 4), 0, stream_, d, r, fl);
            else
              hipLaunchKernelGGL(k_add<8>, dim3(add_grid), dim3(kWave * 8), 0, stream_, d, r, fl);
            if (timing_) HIPCHK(hipEventRecord(ev_[4 * t + 1], stream_));
            hipLaunchKernelGGL(k_add_long, dim3(long_grid), dim3(kLongThreads), 0, stream_, d, r, fl);
            if (timing_) HIPCHK(hipEventRecord(ev_[4 * t + 2], stream_));
            hipLaunchKernelGGL(k_add_huge, dim3(kHugeGroups), dim3(kLongThreads), 0, stream_, d, r, fl);
            if (timing_) HIPCHK(hipEventRecord(ev_[4 * t + 3], stream_));
          }
          HIPCHK(hipGetLastError());
          if ((rc = sync_read(d.ctr, cw::kWords))) return rc;
          S->n_add_launches += rounds_per_sync;
          if (timing_)
            for (int t = 0; t < rounds_per_sync; t++) {
              float ms = 0.f, ms2 = 0.f, ms3 = 0.f;
              HIPCHK(hipEventElapsedTime(&ms, ev_[4 * t], ev_[4 * t + 1]));
              HIPCHK(hipEventElapsedTime(&ms2, ev_[4 * t + 1], ev_[4 * t + 2]));
              HIPCHK(hipEventElapsedTime(&ms3, ev_[4 * t + 2], ev_[4 * t + 3]));
              S->t_add_kernel_ms += ms;
              S->t_add_long_kernel_ms += ms2;
              S->t_add_huge_kernel_ms += ms3;
              if (trace_)
                std::fprintf(stderr, "[ssq] dim %d page %lld round %d: k_add %.3f ms, k_add_long %.3f ms, k_add_huge %.3f ms\n",
                             dim, (long long)S->n_pages, r - rounds_per_sync + t, ms, ms2, ms3);
            }
          if (trace_)
            std::fprintf(stderr,
                         "[ssq] dim %d page %lld after round %d: window %lld, next worklist %u, last long list %u, "
                         "additions so far %llu, longest chain %u, longest |a|+|b| %u\n",
                         dim, (long long)S->n_pages, r - 1, (long long)wn, host_ctr_[cw::kW + (r & 1)],
                         host_ctr_[cw::kLong], ((const unsigned long long*)(host_ctr_ + cw::kStats))[0], host_ctr_[cw::kChain],
                         host_ctr_[cw::kChain + 1]);
          S->n_rounds += rounds_per_sync;
          if (host_ctr_[cw::kErr]) {
            oom = true;
            break;
          }
          if (host_ctr_[cw::kW + (r & 1)] == 0) break;
          if (r > (1 << 24)) {
            err_ = "page did not converge";
            return SSQ_EDEVICE;
          }
        }
      }
      S->n_pages++;
      // a page that covered every column with every row active finished the dimension
      if (end >= nl && mrow <= dminrow) break;
      ss = (int32_t)((float)ss * p.growth);
      if (ss < 1) ss = 1;
      if (xrow <= dminrow) begin = end;
      else {
        mrow -= ss;
        xrow -= ss;
      }
      if (end < nl) end = std::min<int64_t>(end + ss, nl);
    }
    if (oom) break;
    hipLaunchKernelGGL(k_twist, dim3(grid_for(nl)), dim3(256), 0, stream_, d, cols, (int32_t)nl);
  }
  *oom_out = oom;
  return SSQ_OK;
}

int Engine::reduce_device(int64_t n, int64_t nnz, const int64_t* d_colptr, const int32_t* d_rows,
                          const int32_t* d_dims, const ssq_params& p, int64_t* d_pairs, int64_t* np_out,
                          ssq_stats* st, int64_t* d_essential, int64_t* ne_out) {
  if (!np_out || (n > 0 && !d_pairs)) {
    err_ = "invalid arguments";
    return SSQ_EINVAL;
  }
  int rc = init_phase(n, nnz, d_colptr, d_rows, d_dims, p, st);
  if (rc) return rc;
  return compute_phase(d_pairs, np_out, st, d_essential, ne_out);
}

// Phase one (specseq.h:513-590 GPU_Init): validation, buffers, edge collapse, the anti-transpose
// or guided clearing, then the column store, setup, init passes and pending lists.
int Engine::init_phase(int64_t n, int64_t nnz, const int64_t* d_colptr, const int32_t* d_rows, const int32_t* d_dims,
                       const ssq_params& p, ssq_stats* st) {
  err_.clear();
  ph_.ready = false;
  if (n < 0 || nnz < 0 || n >= 0x7fffffffLL || (n > 0 && (!d_colptr || !d_dims))) {
    err_ = "invalid arguments";
    return SSQ_EINVAL;
  }
  if (p.ss_block_size < 1 || !(p.growth >= 1.0f) || p.queue_size < 1 || p.level1 < 1 || p.level2 < 1 ||
      p.max_iteration < 0 || (p.model & ~3) || (p.dualize & ~1) || !shape_ok(p)) {
    err_ = "invalid parameters";
    return SSQ_EINVAL;
  }
  HIPCHK(hipSetDevice(device_));
  std::memset(&ph_.S, 0, sizeof(ph_.S));
  ph_.T0 = std::chrono::steady_clock::now();
  ph_.n = n;
  ph_.p = p;
  if (n == 0) {
    ph_.ready = true;
    if (st) *st = ph_.S;
    return SSQ_OK;
  }
  // (d_colptr, d_rows, d_dims are re-pointed at the dual matrix when p.dualize is set)
  Dev& d = ph_.d;
  d = Dev{};
  d.n = n;
  d.ncols = n;
  d.short_max = short_max_;
  d.long_window = long_window_;
  d.dims = d_dims;
  d.colptr = d_colptr;
  int rc;
  if ((rc = ensure(b_off_, (size_t)n, &d.off)) || (rc = ensure(b_len_, (size_t)n, &d.len)) ||
      (rc = ensure(b_cap_, (size_t)n, &d.cap)) || (rc = ensure(b_low_, (size_t)n, &d.low)) ||
      (rc = ensure(b_state_, (size_t)n, &d.state)) || (rc = ensure(b_removed_, (size_t)n, &d.removed)) ||
      (rc = ensure(b_flag_, (size_t)(2 * n), &d.flag)) || (rc = ensure(b_piv_, (size_t)n, &d.piv)) ||
      (rc = ensure(b_left_, (size_t)n, &d.leftone)) || (rc = ensure(b_pmeta_, (size_t)n, &d.pivmeta)) ||
      (rc = ensure(b_U_, (size_t)n, &ph_.U)))
    return rc;
  d.ctr = ctr_;  // the counter block is the engine's own (read back by the host every few rounds)
  DCHK(dfill(d.ctr, 0, cw::kWords * sizeof(uint32_t)));
  ctr_dev_ = d.ctr;
  d.stats = (unsigned long long*)(d.ctr + cw::kStats);
  d.arena_used = (unsigned long long*)(d.ctr + cw::kArena);
  d.lvl_cnt = d.ctr + cw::kLvl;
  d.err = d.ctr + cw::kErr;
  DCHK(dfill(d.removed, 0, (size_t)n));

  // max dimension (and range check of the cell dimensions)
  hipLaunchKernelGGL(k_maxdim, dim3(grid_for(n)), dim3(256), 0, stream_, n, d_dims, (int32_t*)(d.ctr + cw::kMaxDim),
                     d.ctr + cw::kMaxDim + 1);
  HIPCHK(hipGetLastError());
  if ((rc = sync_read(d.ctr + cw::kMaxDim, 2))) return rc;
  if (host_ctr_[1]) {
    err_ = "cell dimension outside [0, 7]";
    return SSQ_EFORMAT;
  }
  ph_.maxdim = (int32_t)host_ctr_[0];

  // column store: input rows first, then the arena
  ph_.arena_cap = std::max<int64_t>({arena_hint_, 4 * nnz + 4096, (int64_t)1 << 22});
  ssq_stats* S = &ph_.S;
  if ((p.model & 2) && ph_.maxdim >= 1 && (rc = collapse(d, n, d_dims, d_colptr, d_rows, S))) return rc;
  ph_.hd = nullptr;
  ph_.dual = p.dualize == 1;
  ph_.primal_colptr = nullptr;
  ph_.primal_rows = nullptr;
  ph_.primal_removed = nullptr;
  ph_.nnz_eff = nnz;
  ph_.rows = d_rows;
  if (ph_.dual) {
    // cohomology on the (collapsed) complex: reduce the anti-transpose; guided clearing is a
    // statement about the top dimension of the primal complex and is not used here
    int64_t* dcp;
    int32_t *drw, *ddm;
    ph_.primal_colptr = d_colptr;
    ph_.primal_rows = d_rows;
    if ((rc = ensure(b_prem_, (size_t)n, &ph_.primal_removed))) return rc;
    DCHK(dcopy(ph_.primal_removed, d.removed, (size_t)n));
    rc = build_dual(d, n, nnz, ph_.maxdim, d_colptr, d_rows, d_dims, (p.model & 2) != 0, &dcp, &drw, &ddm);
    if (rc) return rc;
    if (p.model & 2) {  // removed cells drop their entries from the dual
      HIPCHK(hipMemcpyAsync(host_tot_ + kTotScratch, dcp + n, sizeof(int64_t), hipMemcpyDeviceToHost, stream_));
      HIPCHK(hipStreamSynchronize(stream_));
      ph_.nnz_eff = host_tot_[kTotScratch];
    }
    ph_.rows = drw;
    d.colptr = dcp;
    d.dims = ddm;
  } else if ((p.model & 1) && ph_.maxdim >= 1) {
    if ((rc = guided_clear(d, n, ph_.maxdim, d_dims, d_colptr, d_rows, p, &ph_.hd, S))) return rc;
  }
  if ((rc = prepare_store())) return rc;
  ph_.T1 = std::chrono::steady_clock::now();
  S->t_init_ms = std::chrono::duration<double, std::milli>(ph_.T1 - ph_.T0).count();
  ph_.ready = true;
  if (st) *st = *S;
  return SSQ_OK;
}

// The column store from the matrix (rows, then an arena of ph_.arena_cap entries), column
// setup, the init passes and the pending lists.  Run again, with a larger arena, when the page
// loop exhausts the store.
int Engine::prepare_store() {
  Dev& d = ph_.d;
  const int64_t n = ph_.n;
  int rc;
  if ((rc = ensure(b_pool_, (size_t)(ph_.nnz_eff + ph_.arena_cap), &d.pool))) return rc;
  d.arena_base = ph_.nnz_eff;
  d.arena_cap = ph_.arena_cap;
  DCHK(dcopy(d.pool, ph_.rows, sizeof(int32_t) * std::max<int64_t>(ph_.nnz_eff, 0)));
  DCHK(dfill(d.ctr, 0, cw::kWords * sizeof(uint32_t)));
  const int g = grid_for(n);
  hipLaunchKernelGGL(k_setup, dim3(g), dim3(256), 0, stream_, d, 1);
  if (ph_.dual)
    hipLaunchKernelGGL(k_dual_init_from_primal, dim3(g), dim3(256), 0, stream_, n, ph_.primal_colptr, ph_.primal_rows,
                       (const uint8_t*)ph_.primal_removed, d.leftone, d.flag + n);
  else
    hipLaunchKernelGGL(k_leftone, dim3(g), dim3(256), 0, stream_, d);
  ph_.L = LoopPlan{};
  ph_.L.U = ph_.U;
  return init_lists(d, ph_.maxdim, (const uint8_t*)ph_.hd, ph_.dual, ph_.p, &ph_.L, &ph_.S);
}

// Phase two (specseq.h:592-1060 GPU_Compute, 1182-1193 GetPersistencePairs): the page loop,
// then the pairs in row order and the essential cells; one round trip at the end.
int Engine::compute_phase(int64_t* d_pairs, int64_t* np_out, ssq_stats* st, int64_t* d_essential, int64_t* ne_out) {
  err_.clear();
  if (!ph_.ready) {
    err_ = "no matrix initialized on this handle";
    return SSQ_EINVAL;
  }
  if (!np_out || (ph_.n > 0 && !d_pairs)) {
    err_ = "invalid arguments";
    return SSQ_EINVAL;
  }
  ph_.ready = false;  // the state is consumed (and changed) by the page loop
  HIPCHK(hipSetDevice(device_));
  *np_out = 0;
  if (ne_out) *ne_out = 0;
  ssq_stats* S = &ph_.S;
  const int64_t n = ph_.n;
  if (n == 0) {
    if (st) *st = *S;
    return SSQ_OK;
  }
  const auto Tc = std::chrono::steady_clock::now();
  Dev& d = ph_.d;
  int rc;
  for (int attempt = 0;; attempt++) {
    bool oom = false;
    if ((rc = page_loop(d, ph_.p, ph_.maxdim, ph_.L, S, &oom))) return rc;
    if (!oom) break;
    const int64_t budget = (int64_t)1 << 36;  // 256 GiB of int32 entries at most
    if (ph_.arena_cap * 2 > budget || attempt > 8) {
      err_ = "column store exhausted";
      return SSQ_ENOMEM;
    }
    ph_.arena_cap *= 2;
    S->n_pages = S->n_rounds = 0;
    if ((rc = prepare_store())) return rc;  // rerun from the matrix with a larger arena
  }
  if (!ph_.dual) {
    if ((rc = compact(n, 1, KeyPivot{d.piv}, EmitPair{d.piv, d_pairs}, nullptr, kTotPairs))) return rc;
  } else {  // dual pair (b', d') -> (n-1-d', n-1-b'), sorted by the new birth
    int32_t* death_of = (int32_t*)d.leftone;
    DCHK(dfill(death_of, 0xff, sizeof(int32_t) * n));
    hipLaunchKernelGGL(k_dual_pairs_fill, dim3(grid_for(n)), dim3(256), 0, stream_, n, d.piv, death_of);
    if ((rc = compact(n, 1, KeyDeathOf{death_of}, EmitDeathOf{death_of, d_pairs}, nullptr, kTotPairs))) return rc;
  }
  if (d_essential && ne_out) {
    int64_t* ess_dst = d_essential;
    if (ph_.dual && (rc = ensure(b_keys_, (size_t)n, &ess_dst))) return rc;
    rc = compact(n, 1, KeyEssential{d.piv, d.removed, d.state, d.low}, EmitI64{ess_dst}, nullptr, kTotEss);
    if (rc) return rc;
    if (ph_.dual)
      hipLaunchKernelGGL(k_dual_essential, dim3(grid_for(n)), dim3(256), 0, stream_, tot_ + kTotEss, ess_dst, n,
                         d_essential);
  }
  if ((rc = sync_all())) return rc;
  *np_out = host_tot_[kTotPairs];
  if (d_essential && ne_out) *ne_out = host_tot_[kTotEss];
  fill_stats(S, ph_.nnz_eff, ph_.arena_cap);
  const auto T2 = std::chrono::steady_clock::now();
  S->t_reduce_ms = std::chrono::duration<double, std::milli>(T2 - Tc).count();
  S->t_total_ms = S->t_init_ms + S->t_reduce_ms;
  if (st) *st = *S;
  return SSQ_OK;
}

// Counters of the last page loop (host_ctr_ must hold the device counters).
void Engine::fill_stats(ssq_stats* S, int64_t nnz, int64_t arena_cap) {
  const unsigned long long* s64 = (const unsigned long long*)(host_ctr_ + cw::kStats);
  S->n_additions = (int64_t)s64[0];
  S->add_bytes = (double)s64[1];
  S->store_bytes = (nnz + arena_cap) * 4;
  arena_hint_ = std::max(arena_hint_, (int64_t)(*(const unsigned long long*)(host_ctr_ + cw::kArena)) * 2);
  S->n_tiled_additions = (int64_t)s64[2];
  S->n_dense_additions = (int64_t)s64[3];
  S->dense_bytes = (double)s64[4];
  S->n_huge_additions = (int64_t)s64[5];
  S->huge_bytes = (double)s64[6];
  S->max_chain = host_ctr_[cw::kChain];
  S->max_addition_len = host_ctr_[cw::kChain + 1];
}

int Engine::prebuilt_begin(int64_t n, int64_t ncols, int64_t nnz, int64_t arena, const int64_t* colptr,
                           const int32_t* dims, int32_t maxdim, PrebuiltBufs* out) {
  err_.clear();
  ph_.ready = false;  // the buffers are reused: a matrix left by init_phase is discarded
  if (n < 1 || ncols < 1 || ncols > n || n >= 0x7fffffffLL || nnz < 0 || arena < 0 || !colptr || !dims || !out ||
      maxdim < 1 || maxdim >= kMaxDims) {
    err_ = "invalid prebuilt matrix";
    return SSQ_EINVAL;
  }
  HIPCHK(hipSetDevice(device_));
  Dev d{};
  d.n = n;
  d.ncols = ncols;
  d.short_max = short_max_;
  d.long_window = long_window_;
  d.dims = dims;
  d.colptr = colptr;
  int rc;
  if ((rc = ensure(b_off_, (size_t)ncols, &d.off)) || (rc = ensure(b_len_, (size_t)ncols, &d.len)) ||
      (rc = ensure(b_cap_, (size_t)ncols, &d.cap)) || (rc = ensure(b_low_, (size_t)ncols, &d.low)) ||
      (rc = ensure(b_state_, (size_t)ncols, &d.state)) || (rc = ensure(b_removed_, (size_t)ncols, &d.removed)) ||
      (rc = ensure(b_flag_, (size_t)(2 * n), &d.flag)) || (rc = ensure(b_piv_, (size_t)n, &d.piv)) ||
      (rc = ensure(b_left_, (size_t)n, &d.leftone)) ||
      (rc = ensure(b_pmeta_, (size_t)n, &d.pivmeta)) || (rc = ensure(b_pool_, (size_t)(nnz + arena), &d.pool)))
    return rc;
  d.arena_base = nnz;
  d.arena_cap = arena;
  d.ctr = ctr_;
  DCHK(dfill(d.removed, 0, (size_t)ncols));
  DCHK(dfill(d.flag, 0, (size_t)(2 * n)));
  pre_ = d;
  pre_maxdim_ = maxdim;
  out->pool = d.pool;
  out->leftone = d.leftone;
  out->negflag = d.flag + n;
  return SSQ_OK;
}

int Engine::prebuilt_reduce(const ssq_params& p, ssq_stats* st, bool* oom) {
  *oom = false;
  if (p.ss_block_size < 1 || !(p.growth >= 1.0f) || p.queue_size < 1 || p.level1 < 1 || p.level2 < 1 ||
      !shape_ok(p)) {
    err_ = "invalid parameters";
    return SSQ_EINVAL;
  }
  HIPCHK(hipSetDevice(device_));
  auto T0 = std::chrono::steady_clock::now();
  ssq_stats local{};
  ssq_stats* S = st ? st : &local;
  std::memset(S, 0, sizeof(*S));
  Dev& d = pre_;
  int32_t* U;
  int rc;
  if ((rc = ensure(b_U_, (size_t)d.ncols, &U))) return rc;
  d.ctr = ctr_;
  DCHK(dfill(d.ctr, 0, cw::kWords * sizeof(uint32_t)));
  ctr_dev_ = d.ctr;
  d.stats = (unsigned long long*)(d.ctr + cw::kStats);
  d.arena_used = (unsigned long long*)(d.ctr + cw::kArena);
  d.lvl_cnt = d.ctr + cw::kLvl;
  d.err = d.ctr + cw::kErr;
  DCHK(dfill(d.piv, 0xff, sizeof(uint32_t) * d.n));
  hipLaunchKernelGGL(k_setup, dim3(grid_for(d.ncols)), dim3(256), 0, stream_, d, 0);
  LoopPlan L;
  L.U = U;
  if ((rc = init_lists(d, pre_maxdim_, nullptr, true, p, &L, S))) return rc;
  auto T1 = std::chrono::steady_clock::now();
  S->t_init_ms = std::chrono::duration<double, std::milli>(T1 - T0).count();
  if ((rc = page_loop(d, p, pre_maxdim_, L, S, oom))) return rc;
  if ((rc = sync_all())) return rc;
  fill_stats(S, d.arena_base, d.arena_cap);
  auto T2 = std::chrono::steady_clock::now();
  S->t_reduce_ms = std::chrono::duration<double, std::milli>(T2 - T1).count();
  S->t_total_ms = std::chrono::duration<double, std::milli>(T2 - T0).count();
  return SSQ_OK;
}

// ====================================================================================
// C ABI
// ====================================================================================
struct PinBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct Handle {
  Engine eng;
  std::string err;
  DevBuf in_colptr, in_rows, in_dims, out_pairs, out_ess;
  PinBuf pin_in, pin_out;
  VrBuilder* vr = nullptr;
  RipsCo* rc = nullptr;
  // further engines and builders of the batched Rips path: its groups run on several streams
  std::vector<std::unique_ptr<Engine>> xeng;
  std::vector<RipsCo*> xrc;
  int rips_streams = 0;  // 0: SSQ_RIPS_STREAMS or 4
  int64_t init_cells = -1;  // cells of the matrix ssq_init left on the engine (-1: none)
  DevBuf pts;
  explicit Handle(int dev) : eng(dev) {}
  ~Handle() {
    if (vr) vr_builder_destroy(vr);
    if (rc) ripsco_destroy(rc);
    for (RipsCo* b : xrc) ripsco_destroy(b);
    if (pts.p) (void)hipFree(pts.p);
    for (DevBuf* b : {&in_colptr, &in_rows, &in_dims, &out_pairs, &out_ess})
      if (b->p) (void)hipFree(b->p);
    for (PinBuf* b : {&pin_in, &pin_out})
      if (b->p) (void)hipHostFree(b->p);
  }
};

// Pinned host staging, grown on demand and kept by the handle.
static int pin_grow(Handle* H, PinBuf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 64);
  if (b.bytes >= bytes) return SSQ_OK;
  if (b.p) (void)hipHostFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  const size_t want = bytes + bytes / 2;
  if (hipHostMalloc(&b.p, want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    H->err = "pinned host allocation failed";
    return SSQ_ENOMEM;
  }
  b.bytes = want;
  return SSQ_OK;
}

static int dev_grow(Handle* H, DevBuf& b, size_t bytes) {
  bytes = std::max<size_t>(bytes, 64);
  if (b.bytes >= bytes) return SSQ_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes + bytes / 2) != hipSuccess) {
    (void)hipGetLastError();
    H->err = "device allocation failed";
    return SSQ_ENOMEM;
  }
  b.bytes = bytes + bytes / 2;
  return SSQ_OK;
}

// memcpy of many pieces; large totals are split over host threads in 1 MiB chunks
struct CopyJob {
  void* dst;
  const void* src;
  size_t bytes;
};
static void par_copy(const std::vector<CopyJob>& jobs) {
  size_t total = 0;
  for (const CopyJob& j : jobs) total += j.bytes;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int threads = total < (size_t(8) << 20) ? 1 : (int)std::min<unsigned>(8, hw);
  if (threads == 1) {
    for (const CopyJob& j : jobs)
      if (j.bytes) std::memcpy(j.dst, j.src, j.bytes);
    return;
  }
  constexpr size_t kChunk = size_t(1) << 20;
  std::vector<CopyJob> pieces;
  for (const CopyJob& j : jobs)
    for (size_t o = 0; o < j.bytes; o += kChunk)
      pieces.push_back({(char*)j.dst + o, (const char*)j.src + o, std::min(kChunk, j.bytes - o)});
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (size_t k; (k = next.fetch_add(1)) < pieces.size();) std::memcpy(pieces[k].dst, pieces[k].src, pieces[k].bytes);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; t++) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
}

static int split_pairs(Handle* H, const int64_t* dp, int64_t np, const int64_t* de, int64_t ne, const int64_t* base,
                       int count, ssq_pairs* outs);

// Reduce `count` host matrices as one block-diagonal batch (count == 1: a single matrix).
// Host work is O(count) checks plus one staged copy; the per-nonzero validation runs on the
// device before any kernel indexes with the input.  outs[i] receives block i's pairs in its
// own indices.
// The matrices into the handle's device buffers, block-diagonally (checked on the device:
// rows in range and strictly increasing per column), shifted to global indices.
struct Uploaded {
  std::vector<int64_t> base;  // first cell of each block, then the total
  int64_t N = 0, NNZ = 0;
  int64_t* colptr = nullptr;
  int32_t* rows = nullptr;
  int32_t* dims = nullptr;
};
static int upload_blocks(Handle* H, const ssq_matrix* const* mats, int count, Uploaded* U) {
  std::vector<int64_t> base((size_t)count + 1, 0), nzb((size_t)count + 1, 0);
  for (int i = 0; i < count; i++) {
    const ssq_matrix* m = mats[i];
    if (!m || m->n < 0 || m->nnz < 0 || (m->n > 0 && (!m->col_ptr || !m->dims)) || (m->nnz > 0 && !m->rows)) {
      H->err = "invalid matrix " + std::to_string(i);
      return SSQ_EINVAL;
    }
    if (m->n > 0 && (m->col_ptr[0] != 0 || m->col_ptr[m->n] != m->nnz)) {
      H->err = "malformed matrix " + std::to_string(i);
      return SSQ_EFORMAT;
    }
    if (m->n == 0 && m->nnz != 0) {
      H->err = "malformed matrix " + std::to_string(i);
      return SSQ_EFORMAT;
    }
    base[i + 1] = base[i] + m->n;
    nzb[i + 1] = nzb[i] + m->nnz;
  }
  const int64_t N = base[count], NNZ = nzb[count];
  if (N >= 0x7fffffffLL) {
    H->err = "more than 2^31 - 1 cells";
    return SSQ_EINVAL;
  }
  // staging layout: colptr[N + 1] | base[count + 1] | nzb[count + 1] | rows[NNZ] | dims[N]
  const size_t o_base = sizeof(int64_t) * (N + 1), o_nzb = o_base + sizeof(int64_t) * (count + 1),
               o_rows = o_nzb + sizeof(int64_t) * (count + 1), o_dims = o_rows + sizeof(int32_t) * NNZ,
               total = o_dims + sizeof(int32_t) * N;
  int rc;
  if ((rc = pin_grow(H, H->pin_in, total))) return rc;
  char* stg = (char*)H->pin_in.p;
  std::vector<CopyJob> jobs;
  for (int i = 0; i < count; i++) {
    const ssq_matrix* m = mats[i];
    jobs.push_back({stg + sizeof(int64_t) * base[i], m->col_ptr, sizeof(int64_t) * m->n});
    jobs.push_back({stg + o_rows + sizeof(int32_t) * nzb[i], m->rows, sizeof(int32_t) * m->nnz});
    jobs.push_back({stg + o_dims + sizeof(int32_t) * base[i], m->dims, sizeof(int32_t) * m->n});
  }
  par_copy(jobs);
  ((int64_t*)stg)[N] = NNZ;  // the shift below leaves it in place for the last block
  std::memcpy(stg + o_base, base.data(), sizeof(int64_t) * (count + 1));
  std::memcpy(stg + o_nzb, nzb.data(), sizeof(int64_t) * (count + 1));
  if ((rc = dev_grow(H, H->in_colptr, o_rows)) || (rc = dev_grow(H, H->in_rows, sizeof(int32_t) * NNZ)) ||
      (rc = dev_grow(H, H->in_dims, sizeof(int32_t) * N + 64)) ||
      (rc = dev_grow(H, H->out_pairs, sizeof(int64_t) * 2 * N)) || (rc = dev_grow(H, H->out_ess, sizeof(int64_t) * N)))
    return rc;
  hipStream_t s = H->eng.stream();
  int64_t* dc = (int64_t*)H->in_colptr.p;  // colptr, then base and nzb
  int64_t* dbase = dc + N + 1;
  int64_t* dnzb = dbase + count + 1;
  int32_t* dr = (int32_t*)H->in_rows.p;
  int32_t* dd = (int32_t*)H->in_dims.p;
  uint32_t* dbad = (uint32_t*)(dd + N + ((N & 1) ? 1 : 0));  // 8-byte aligned word after the dims
  if (hipMemcpyAsync(dc, stg, o_rows, hipMemcpyHostToDevice, s) != hipSuccess ||
      (NNZ && hipMemcpyAsync(dr, stg + o_rows, sizeof(int32_t) * NNZ, hipMemcpyHostToDevice, s) != hipSuccess) ||
      (N && hipMemcpyAsync(dd, stg + o_dims, sizeof(int32_t) * N, hipMemcpyHostToDevice, s) != hipSuccess) ||
      hipMemsetAsync(dbad, 0, sizeof(uint32_t), s) != hipSuccess) {
    H->err = "host to device copy failed";
    return SSQ_EDEVICE;
  }
  if (N > 0) {
    const int grid = (int)std::max<int64_t>(8, std::min<int64_t>((N + 255) / 256, 4096) / 8 * 8);
    hipLaunchKernelGGL(k_check_blocks, dim3(grid), dim3(256), 0, s, N, (int64_t)count, dbase, dnzb, dc, dr, dbad);
    uint32_t bad = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&bad, dbad, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      H->err = "device validation failed";
      return SSQ_EDEVICE;
    }
    if (bad) {
      H->err = count > 1 ? "malformed matrix in batch (row out of range or rows not strictly increasing)"
                         : "malformed matrix (row out of range or rows not strictly increasing)";
      return SSQ_EFORMAT;
    }
    if (count > 1) {
      const int64_t work = std::max<int64_t>(N, NNZ);
      const int grid2 = (int)std::max<int64_t>(8, std::min<int64_t>((work + 255) / 256, 4096) / 8 * 8);
      hipLaunchKernelGGL(k_batch_shift, dim3(grid2), dim3(256), 0, s, N, NNZ, (int64_t)count, dbase, dnzb, dc, dr);
      if (hipGetLastError() != hipSuccess) return SSQ_EDEVICE;
    }
  }
  U->base = std::move(base);
  U->N = N;
  U->NNZ = NNZ;
  U->colptr = dc;
  U->rows = dr;
  U->dims = dd;
  return SSQ_OK;
}

// Matrices already resident in HBM (validated when they were placed there) into the batch
// buffers: one gather kernel per block, no host staging, no host round trip.
static int gather_blocks(Handle* H, const ssq_device_matrix* mats, int count, Uploaded* U) {
  std::vector<int64_t> base((size_t)count + 1, 0), nzb((size_t)count + 1, 0);
  for (int i = 0; i < count; i++) {
    const ssq_device_matrix& m = mats[i];
    if (m.n < 0 || m.nnz < 0 || (m.n > 0 && (!m.col_ptr || !m.dims)) || (m.nnz > 0 && !m.rows) ||
        (m.n == 0 && m.nnz != 0)) {
      H->err = "invalid resident matrix " + std::to_string(i);
      return SSQ_EINVAL;
    }
    base[i + 1] = base[i] + m.n;
    nzb[i + 1] = nzb[i] + m.nnz;
  }
  const int64_t N = base[count], NNZ = nzb[count];
  if (N >= 0x7fffffffLL) {
    H->err = "more than 2^31 - 1 cells";
    return SSQ_EINVAL;
  }
  int rc;
  if ((rc = dev_grow(H, H->in_colptr, sizeof(int64_t) * (N + 1))) ||
      (rc = dev_grow(H, H->in_rows, sizeof(int32_t) * NNZ)) || (rc = dev_grow(H, H->in_dims, sizeof(int32_t) * N)) ||
      (rc = dev_grow(H, H->out_pairs, sizeof(int64_t) * 2 * N)) || (rc = dev_grow(H, H->out_ess, sizeof(int64_t) * N)))
    return rc;
  hipStream_t s = H->eng.stream();
  int64_t* dc = (int64_t*)H->in_colptr.p;
  int32_t* dr = (int32_t*)H->in_rows.p;
  int32_t* dd = (int32_t*)H->in_dims.p;
  if (N == 0 && (rc = H->eng.fill(dc, 0, sizeof(int64_t)))) return rc;  // colptr[0] = 0
  for (int i = 0; i < count; i++) {
    const ssq_device_matrix& m = mats[i];
    if (m.n == 0) continue;
    const int64_t work = std::max<int64_t>(m.n + 1, m.nnz);
    const int grid = (int)std::max<int64_t>(8, std::min<int64_t>((work + 255) / 256, 2048) / 8 * 8);
    hipLaunchKernelGGL(k_gather_block, dim3(grid), dim3(256), 0, s, m.n, m.nnz, m.col_ptr, m.rows, m.dims, base[i],
                       nzb[i], dc + base[i], dr + nzb[i], dd + base[i]);
    if (hipGetLastError() != hipSuccess) {
      H->err = "gather launch failed";
      return SSQ_EDEVICE;
    }
  }
  U->base = std::move(base);
  U->N = N;
  U->NNZ = NNZ;
  U->colptr = dc;
  U->rows = dr;
  U->dims = dd;
  return SSQ_OK;
}

static int reduce_blocks(Handle* H, const ssq_matrix* const* mats, int count, const ssq_params& p, ssq_pairs* outs,
                         ssq_stats* st) {
  for (int i = 0; i < count; i++) std::memset(&outs[i], 0, sizeof(ssq_pairs));
  Uploaded U;
  int rc = upload_blocks(H, mats, count, &U);
  if (rc) return rc;
  int64_t np = 0, ne = 0;
  int64_t* dp = (int64_t*)H->out_pairs.p;
  int64_t* de = (int64_t*)H->out_ess.p;
  rc = H->eng.reduce_device(U.N, U.NNZ, U.colptr, U.rows, U.dims, p, dp, &np, st, de, &ne);
  if (rc) {
    H->err = H->eng.error();
    return rc;
  }
  return split_pairs(H, dp, np, de, ne, U.base.data(), count, outs);
}

// Copy the batch's pairs and essentials to the host and give each block its own run, in its
// own indices (pairs are sorted by birth and essentials ascending).
static int split_pairs(Handle* H, const int64_t* dp, int64_t np, const int64_t* de, int64_t ne, const int64_t* base,
                       int count, ssq_pairs* outs) {
  hipStream_t s = H->eng.stream();
  int rc;
  if ((rc = pin_grow(H, H->pin_out, sizeof(int64_t) * (2 * np + ne)))) return rc;
  int64_t* hp = (int64_t*)H->pin_out.p;
  int64_t* he = hp + 2 * np;
  if ((np && hipMemcpyAsync(hp, dp, sizeof(int64_t) * 2 * np, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (ne && hipMemcpyAsync(he, de, sizeof(int64_t) * ne, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    H->err = "device to host copy failed";
    return SSQ_EDEVICE;
  }
  int64_t pi = 0, ei = 0;
  for (int i = 0; i < count; i++) {
    int64_t lo = pi, hi = np;  // first pair with birth >= base[i + 1]
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (hp[2 * mid] < base[i + 1]) lo = mid + 1;
      else hi = mid;
    }
    const int64_t pj = lo;
    int64_t elo = ei, ehi = ne;
    while (elo < ehi) {
      const int64_t mid = (elo + ehi) >> 1;
      if (he[mid] < base[i + 1]) elo = mid + 1;
      else ehi = mid;
    }
    const int64_t ej = elo;
    ssq_pairs* o = &outs[i];
    o->np = pj - pi;
    o->ne = ej - ei;
    o->pairs = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(2 * o->np, 1));
    o->essential = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(o->ne, 1));
    if (!o->pairs || !o->essential) {
      for (int k = 0; k <= i; k++) ssq_pairs_free(&outs[k]);
      H->err = "host allocation failed";
      return SSQ_ENOMEM;
    }
    const int64_t b = base[i];
    for (int64_t k = 0; k < 2 * o->np; k++) o->pairs[k] = hp[2 * pi + k] - b;
    for (int64_t k = 0; k < o->ne; k++) o->essential[k] = he[ei + k] - b;
    pi = pj;
    ei = ej;
  }
  return SSQ_OK;
}

}  // namespace ssq

using namespace ssq;

struct ssq_handle {
  Handle* h;
};

extern "C" int ssq_create(int device, ssq_handle** out) {
  if (!out) return SSQ_EINVAL;
  *out = nullptr;
  Handle* H = new (std::nothrow) Handle(device);
  if (!H) return SSQ_ENOMEM;
  int rc = H->eng.init();
  if (rc) {
    delete H;
    return rc;
  }
  ssq_handle* h = new (std::nothrow) ssq_handle{H};
  if (!h) {
    delete H;
    return SSQ_ENOMEM;
  }
  *out = h;
  return SSQ_OK;
}

extern "C" int ssq_destroy(ssq_handle* h) {
  if (!h) return SSQ_EINVAL;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) (void)hipGetLastError();
  delete h->h;
  delete h;
  return SSQ_OK;
}

extern "C" const char* ssq_last_error(const ssq_handle* h) { return h ? h->h->err.c_str() : "null handle"; }

extern "C" int ssq_reduce(ssq_handle* h, const ssq_matrix* m, const ssq_params* p, ssq_pairs* out,
                          ssq_stats* stats) {
  if (!h || !p || !out) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  h->h->err.clear();
  if (!m) return SSQ_EINVAL;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  const ssq_matrix* one[1] = {m};
  return reduce_blocks(h->h, one, 1, *p, out, stats);
}

extern "C" int ssq_init(ssq_handle* h, const ssq_matrix* m, const ssq_params* p, ssq_stats* stats) {
  if (!h || !p || !m) return SSQ_EINVAL;
  Handle* H = h->h;
  H->err.clear();
  H->init_cells = -1;
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  const ssq_matrix* one[1] = {m};
  Uploaded U;
  int rc = upload_blocks(H, one, 1, &U);
  if (rc) return rc;
  if ((rc = H->eng.init_phase(U.N, U.NNZ, U.colptr, U.rows, U.dims, *p, stats))) {
    H->err = H->eng.error();
    return rc;
  }
  H->init_cells = U.N;
  return SSQ_OK;
}

extern "C" int ssq_compute(ssq_handle* h, ssq_pairs* out, ssq_stats* stats) {
  if (!h || !out) return SSQ_EINVAL;
  Handle* H = h->h;
  std::memset(out, 0, sizeof(*out));
  H->err.clear();
  if (H->init_cells < 0) {
    H->err = "no matrix initialized on this handle (ssq_init first)";
    return SSQ_EINVAL;
  }
  const int64_t base[2] = {0, H->init_cells};
  H->init_cells = -1;
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  int64_t np = 0, ne = 0;
  int64_t* dp = (int64_t*)H->out_pairs.p;
  int64_t* de = (int64_t*)H->out_ess.p;
  int rc = H->eng.compute_phase(dp, &np, stats, de, &ne);
  if (rc) {
    H->err = H->eng.error();
    return rc;
  }
  return split_pairs(H, dp, np, de, ne, base, 1, out);
}

extern "C" int ssq_reduce_device(ssq_handle* h, int64_t n, int64_t nnz, const int64_t* d_col_ptr,
                                 const int32_t* d_rows, const int32_t* d_dims, const ssq_params* p,
                                 int64_t* d_pairs_out, int64_t* np_out, ssq_stats* stats) {
  if (!h || !p) return SSQ_EINVAL;
  h->h->err.clear();
  int rc = h->h->eng.reduce_device(n, nnz, d_col_ptr, d_rows, d_dims, *p, d_pairs_out, np_out, stats);
  if (rc) h->h->err = h->h->eng.error();
  return rc;
}

extern "C" int ssq_batch_reduce(ssq_handle* h, const ssq_matrix* const* mats, int count, const ssq_params* p,
                                ssq_pairs* outs, ssq_stats* stats) {
  if (!h || !p || !outs || count < 0 || (count > 0 && !mats)) return SSQ_EINVAL;
  h->h->err.clear();
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  return reduce_blocks(h->h, mats, count, *p, outs, stats);
}

extern "C" int ssq_matrix_to_device(ssq_handle* h, const ssq_matrix* m, ssq_device_matrix* out) {
  if (!h || !m || !out) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  Handle* H = h->h;
  H->err.clear();
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  // through the handle's input buffers, where the matrix is checked on the device
  const ssq_matrix* one[1] = {m};
  Uploaded U;
  int rc = upload_blocks(H, one, 1, &U);
  if (rc) return rc;
  void *cp = nullptr, *rw = nullptr, *dm = nullptr;
  if (hipMalloc(&cp, sizeof(int64_t) * (U.N + 1)) != hipSuccess ||
      hipMalloc(&rw, sizeof(int32_t) * std::max<int64_t>(U.NNZ, 1)) != hipSuccess ||
      hipMalloc(&dm, sizeof(int32_t) * std::max<int64_t>(U.N, 1)) != hipSuccess) {
    (void)hipGetLastError();
    if (cp) (void)hipFree(cp);
    if (rw) (void)hipFree(rw);
    H->err = "device allocation failed";
    return SSQ_ENOMEM;
  }
  if ((rc = H->eng.copy(cp, U.colptr, sizeof(int64_t) * (U.N + 1))) ||
      (rc = H->eng.copy(rw, U.rows, sizeof(int32_t) * U.NNZ)) ||
      (rc = H->eng.copy(dm, U.dims, sizeof(int32_t) * U.N)) || hipStreamSynchronize(H->eng.stream()) != hipSuccess) {
    (void)hipFree(cp);
    (void)hipFree(rw);
    (void)hipFree(dm);
    H->err = "device copy failed";
    return rc ? rc : SSQ_EDEVICE;
  }
  out->n = U.N;
  out->nnz = U.NNZ;
  out->col_ptr = (const int64_t*)cp;
  out->rows = (const int32_t*)rw;
  out->dims = (const int32_t*)dm;
  return SSQ_OK;
}

extern "C" int ssq_device_matrix_free(ssq_handle* h, ssq_device_matrix* m) {
  if (!h || !m) return SSQ_EINVAL;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  for (const void* p : {(const void*)m->col_ptr, (const void*)m->rows, (const void*)m->dims})
    if (p) (void)hipFree(const_cast<void*>(p));
  std::memset(m, 0, sizeof(*m));
  return SSQ_OK;
}

extern "C" int ssq_batch_reduce_resident(ssq_handle* h, const ssq_device_matrix* mats, int count, const ssq_params* p,
                                         ssq_pairs* outs, ssq_stats* stats) {
  if (!h || !p || !outs || count < 0 || (count > 0 && !mats)) return SSQ_EINVAL;
  Handle* H = h->h;
  H->err.clear();
  for (int i = 0; i < count; i++) std::memset(&outs[i], 0, sizeof(ssq_pairs));
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  Uploaded U;
  int rc = gather_blocks(H, mats, count, &U);
  if (rc) return rc;
  int64_t np = 0, ne = 0;
  int64_t* dp = (int64_t*)H->out_pairs.p;
  int64_t* de = (int64_t*)H->out_ess.p;
  if ((rc = H->eng.reduce_device(U.N, U.NNZ, U.colptr, U.rows, U.dims, *p, dp, &np, stats, de, &ne))) {
    H->err = H->eng.error();
    return rc;
  }
  return split_pairs(H, dp, np, de, ne, U.base.data(), count, outs);
}

extern "C" int ssq_device_alloc(ssq_handle* h, size_t bytes, void** dptr) {
  if (!h || !dptr) return SSQ_EINVAL;
  *dptr = nullptr;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  if (hipMalloc(dptr, std::max<size_t>(bytes, 1)) != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    h->h->err = "device allocation failed";
    return SSQ_ENOMEM;
  }
  return SSQ_OK;
}

extern "C" int ssq_device_free(ssq_handle* h, void* dptr) {
  if (!h) return SSQ_EINVAL;
  if (dptr && hipFree(dptr) != hipSuccess) return SSQ_EDEVICE;
  return SSQ_OK;
}

static int copy_sync(ssq_handle* h, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (!h || (bytes && (!dst || !src))) return SSQ_EINVAL;
  if (!bytes) return SSQ_OK;
  hipStream_t s = h->h->eng.stream();
  if (hipMemcpyAsync(dst, src, bytes, kind, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    h->h->err = "copy failed";
    return SSQ_EDEVICE;
  }
  return SSQ_OK;
}

extern "C" int ssq_copy_to_device(ssq_handle* h, void* dst, const void* src, size_t bytes) {
  return copy_sync(h, dst, src, bytes, hipMemcpyHostToDevice);
}

extern "C" int ssq_copy_to_host(ssq_handle* h, void* dst, const void* src, size_t bytes) {
  return copy_sync(h, dst, src, bytes, hipMemcpyDeviceToHost);
}

extern "C" int ssq_device_synchronize(ssq_handle* h) {
  if (!h) return SSQ_EINVAL;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return SSQ_EDEVICE;
  return SSQ_OK;
}

// Streams of the batched Rips path: the handle's setting (ssq::handle_set_rips_streams), else
// SSQ_RIPS_STREAMS, else one per hardware queue: HIP spreads a process's streams over
// GPU_MAX_HW_QUEUES queues (4 unless set), and two streams on one queue run one after the
// other.  At most kMaxRipsStreams.  Each stream reduces its own groups; a group's reduction
// keeps few columns chasing at a time, so more groups in flight fill more of the GPU (with
// the budget split over more streams, each group is smaller).
constexpr int kMaxRipsStreams = 16;  // with 4 GiB of device region each, 24 ran out of memory
static int rips_streams(const Handle* H) {
  int n = 4;
  if (const char* q = std::getenv("GPU_MAX_HW_QUEUES")) n = std::atoi(q);
  if (const char* e = std::getenv("SSQ_RIPS_STREAMS")) n = std::atoi(e);
  if (H->rips_streams > 0) n = H->rips_streams;
  return std::max(1, std::min(kMaxRipsStreams, n));
}

static int rips_reduce(Handle* H, int count, const float* d_pts, const int64_t* pbase, int pdim, const float* radius,
                       const ssq_params& p, int64_t* d_pairs, int64_t pairs_cap, int64_t* pair_off, ssq_stats* stats,
                       const int64_t** d_pairs_owned) {
  if (!H->rc && !(H->rc = ripsco_create())) return SSQ_ENOMEM;
  const int ns = rips_streams(H);
  while ((int)H->xeng.size() < ns - 1) {
    std::unique_ptr<Engine> e(new (std::nothrow) Engine(H->eng.device()));
    RipsCo* b = ripsco_create();
    if (!e || !b) {
      if (b) ripsco_destroy(b);
      return SSQ_ENOMEM;
    }
    if (int rc = e->init()) {
      ripsco_destroy(b);
      H->err = e->error();
      return rc;
    }
    H->xeng.push_back(std::move(e));
    H->xrc.push_back(b);
  }
  std::vector<RipsCo*> bs{H->rc};
  std::vector<PrebuiltTarget*> es{&H->eng};
  for (int k = 0; k < ns - 1; k++) {
    bs.push_back(H->xrc[(size_t)k]);
    es.push_back(H->xeng[(size_t)k].get());
  }
  return ripsco_reduce(bs.data(), es.data(), ns, count, d_pts, pbase, pdim, radius, p, d_pairs, pairs_cap, pair_off, 0,
                       stats, &H->err, d_pairs_owned);
}

// Point clouds through the batched Rips coboundary builder: the points go to the device,
// the pairs come back per cloud, and the essential cells are the complement of the pairs in
// each cloud's n + edges + triangles cells.
static int points_cohomology(Handle* H, int count, const float* const* pts, const int64_t* npts, int pdim,
                             const float* r, const ssq_params& p, ssq_pairs* outs, ssq_stats* stats) {
  std::vector<int64_t> pbase((size_t)count + 1, 0);
  for (int i = 0; i < count; i++) {
    if (npts[i] < 0 || (npts[i] && !pts[i]) || !(r[i] >= 0.f)) return SSQ_EINVAL;
    pbase[i + 1] = pbase[i] + npts[i];
  }
  const int64_t np_all = pbase[count];
  int rc;
  if ((rc = pin_grow(H, H->pin_in, sizeof(float) * std::max<int64_t>(np_all * pdim, 1)))) return rc;
  float* stg = (float*)H->pin_in.p;
  for (int i = 0; i < count; i++)
    if (npts[i]) std::memcpy(stg + pbase[i] * pdim, pts[i], sizeof(float) * npts[i] * pdim);
  if ((rc = dev_grow(H, H->pts, sizeof(float) * std::max<int64_t>(np_all * pdim, 1)))) return rc;
  hipStream_t s = H->eng.stream();
  if (np_all && hipMemcpyAsync(H->pts.p, stg, sizeof(float) * np_all * pdim, hipMemcpyHostToDevice, s) != hipSuccess) {
    H->err = "host to device copy failed";
    return SSQ_EDEVICE;
  }
  std::vector<int64_t> off((size_t)count + 1, 0);
  const int64_t* dp = nullptr;
  rc = rips_reduce(H, count, (const float*)H->pts.p, pbase.data(), pdim, r, p, nullptr, 0, off.data(), stats, &dp);
  if (rc) return rc;
  const int64_t np = off[count];
  if ((rc = pin_grow(H, H->pin_out, sizeof(int64_t) * std::max<int64_t>(2 * np, 1)))) return rc;
  int64_t* hp = (int64_t*)H->pin_out.p;
  if ((np && hipMemcpyAsync(hp, dp, sizeof(int64_t) * 2 * np, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess) {
    H->err = "device to host copy failed";
    return SSQ_EDEVICE;
  }
  const int64_t* tris = ripsco_triangles(H->rc);
  const int64_t* edges = ripsco_edges(H->rc);
  for (int i = 0; i < count; i++) {
    ssq_pairs* o = &outs[i];
    const int64_t cells = npts[i] + edges[i] + tris[i];
    o->np = off[i + 1] - off[i];
    o->ne = cells - 2 * o->np;
    o->pairs = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(2 * o->np, 1));
    o->essential = (int64_t*)std::malloc(sizeof(int64_t) * std::max<int64_t>(o->ne, 1));
    std::vector<uint8_t> paired((size_t)cells, 0);
    if (!o->pairs || !o->essential || o->ne < 0) {
      for (int k = 0; k <= i; k++) ssq_pairs_free(&outs[k]);
      H->err = o->ne < 0 ? "pair count exceeds the cell count" : "host allocation failed";
      return o->ne < 0 ? SSQ_EDEVICE : SSQ_ENOMEM;
    }
    std::memcpy(o->pairs, hp + 2 * off[i], sizeof(int64_t) * 2 * o->np);
    for (int64_t k = 0; k < 2 * o->np; k++) {
      const int64_t c = o->pairs[k];
      if (c >= 0 && c < cells) paired[(size_t)c] = 1;
    }
    int64_t e = 0;
    for (int64_t c = 0; c < cells && e < o->ne; c++)
      if (!paired[(size_t)c]) o->essential[e++] = c;
    if (e != o->ne) {
      for (int k = 0; k <= i; k++) ssq_pairs_free(&outs[k]);
      H->err = "pairs repeat a cell";
      return SSQ_EDEVICE;
    }
  }
  return SSQ_OK;
}

extern "C" int ssq_batch_reduce_points(ssq_handle* h, int count, const float* const* pts, const int64_t* npts,
                                       int pdim, const float* r, int maxdim, const ssq_params* p, ssq_pairs* outs,
                                       ssq_stats* stats) {
  if (!h || !p || !outs || count < 1 || !pts || !npts || !r || pdim < 1 || maxdim < 0 || maxdim > 2)
    return SSQ_EINVAL;
  for (int i = 0; i < count; i++) std::memset(&outs[i], 0, sizeof(ssq_pairs));
  Handle* H = h->h;
  H->err.clear();
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  if (maxdim == 2 && p->dualize == 1 && !(p->model & 2) && !std::getenv("SSQ_RIPS_GENERAL")) {
    bool small = true;
    for (int i = 0; i < count && small; i++) small = npts[i] >= 0 && npts[i] <= kRipsMaxPoints;
    if (small) return points_cohomology(H, count, pts, npts, pdim, r, *p, outs, stats);
  }
  if (!H->vr && !(H->vr = vr_builder_create())) return SSQ_ENOMEM;
  std::vector<int64_t> base((size_t)count + 1);
  const int64_t* dc;
  const int32_t *dr, *dd;
  int64_t n = 0, nnz = 0;
  int rc = vr_builder_build_batch(H->vr, H->eng.stream(), count, pts, npts, pdim, r, maxdim, &dc, &dr, &dd, &n, &nnz,
                                  base.data());
  if (rc) {
    H->err = rc == SSQ_EINVAL ? "invalid point clouds" : "GPU Rips build failed";
    return rc;
  }
  if ((rc = dev_grow(H, H->out_pairs, sizeof(int64_t) * 2 * n)) || (rc = dev_grow(H, H->out_ess, sizeof(int64_t) * n)))
    return rc;
  int64_t np = 0, ne = 0;
  int64_t* dp = (int64_t*)H->out_pairs.p;
  int64_t* de = (int64_t*)H->out_ess.p;
  rc = H->eng.reduce_device(n, nnz, dc, dr, dd, *p, dp, &np, stats, de, &ne);
  if (rc) {
    H->err = H->eng.error();
    return rc;
  }
  return split_pairs(H, dp, np, de, ne, base.data(), count, outs);
}

extern "C" int ssq_vr_build_gpu(ssq_handle* h, const float* pts, int64_t npts, int pdim, float r, int maxdim,
                                ssq_matrix* out) {
  if (!h || !out) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  if (!h->h->vr && !(h->h->vr = vr_builder_create())) return SSQ_ENOMEM;
  if (hipSetDevice(h->h->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  const int64_t* dc;
  const int32_t *dr, *dd;
  int64_t n = 0, nnz = 0;
  int rc = vr_builder_build(h->h->vr, h->h->eng.stream(), pts, npts, pdim, r, maxdim, &dc, &dr, &dd, &n, &nnz);
  if (rc) {
    h->h->err = "GPU Rips build failed";
    return rc;
  }
  std::vector<int64_t> cp((size_t)n + 1);
  std::vector<int32_t> rw((size_t)nnz), dm((size_t)n);
  hipStream_t s = h->h->eng.stream();
  if (hipMemcpyAsync(cp.data(), dc, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
      (nnz && hipMemcpyAsync(rw.data(), dr, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      (n && hipMemcpyAsync(dm.data(), dd, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess) ||
      hipStreamSynchronize(s) != hipSuccess)
    return SSQ_EDEVICE;
  return matrix_from_vectors(cp, rw, dm, out);
}

extern "C" int ssq_rips_count(ssq_handle* h, int count, const float* d_pts, const int64_t* pbase, int pdim,
                              const float* radius, int64_t* edges, int64_t* triangles) {
  if (!h || count < 1 || !pbase || !radius || pdim < 1) return SSQ_EINVAL;
  Handle* H = h->h;
  H->err.clear();
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  if (!H->rc && !(H->rc = ripsco_create())) return SSQ_ENOMEM;
  const int rc = ripsco_count(H->rc, H->eng.stream(), count, d_pts, pbase, pdim, radius, edges, triangles);
  if (rc) H->err = "Rips count failed";
  return rc;
}

extern "C" int ssq_batch_reduce_points_device(ssq_handle* h, int count, const float* d_pts, const int64_t* pbase,
                                              int pdim, const float* radius, int maxdim, const ssq_params* p,
                                              int64_t* d_pairs, int64_t pairs_cap, int64_t* pair_off,
                                              ssq_stats* stats) {
  if (!h || !p || count < 1 || !pbase || !radius || pdim < 1 || maxdim != 2 || !d_pairs || !pair_off ||
      (p->model & 2))
    return SSQ_EINVAL;
  Handle* H = h->h;
  H->err.clear();
  if (hipSetDevice(H->eng.device()) != hipSuccess) return SSQ_EDEVICE;
  return rips_reduce(H, count, d_pts, pbase, pdim, radius, *p, d_pairs, pairs_cap, pair_off, stats, nullptr);
}

namespace ssq {
void handle_set_rips_streams(ssq_handle* h, int n) {
  if (h) h->h->rips_streams = n;
}
}  // namespace ssq
