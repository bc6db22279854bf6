// Batched Vietoris-Rips persistence in cohomology (dimension <= 2), built straight into the
// reduction engine's dual layout: points in, pairs out, no primal matrix and no sort of the
// triangles or of the nonzeros.
//
// The complex and its cell order are the host builder's (ssq_gen.cpp; semantics of
// phat-highclear/addons/rips.cpp:49-87): an edge joins points at float32 distance <= r,
// triangles are the 3-cliques, a cell's filtration is its longest edge, and cells are ordered
// by (filtration, dimension, lexicographic vertex ids).  The pairs are those of the twist
// reduction of that matrix; here they come from the coboundary (dual) matrix, reduced by the
// engine's page loop (specseq.h:303-363,598-704 reduce the anti-transpose when dualize is set).
//
// Cell layout handed to the engine, per group of clouds: [vertices | edges | triangles], each
// block cloud-major and, within a cloud, in reverse filtration order (the dual order).  Only
// vertices and edges have (coboundary) columns; triangles are rows only.  Pairs do not depend
// on the relative order of cells of different dimensions (a column's rows all have one
// dimension and additions combine columns of one dimension), which is the reference's own
// reorder by dimension (specseq.h:439-483).
//
// Triangle ranks without a sort.  The edges of a cloud are ranked by (filtration, lex).  For
// the edge m = (u, v) of rank k, Q(m) is the set of vertices x with both (u, x) and (v, x)
// of rank < k: the triangles whose latest edge is m.  If no other edge of the cloud has m's
// length, the triangles of that filtration value are exactly those of Q(m), and their
// (filtration, lex) order is the order of x (for fixed u < v the sorted triple grows with x).
// So triangle {u, v, x} has rank TB[m] + #{y in Q(m) : y < x}, with TB the prefix sum of |Q|
// over the ranked edges.  Runs of equal-length edges get their triangles' order from a sort of
// lexicographic keys over the run (rare outside lattices).
//
// Coboundary of edge e = (a, b): the triangles {a, b, x} over the common neighbours x.  Those
// whose latest edge is e itself ("own") have the smallest ranks, contiguous from TB[e].  Any
// other has its latest edge m = (a, x) or (b, x), which shares exactly one endpoint with e;
// that entry is computed by the workgroup of the shared endpoint, which stages the Q bitsets
// of all its edges in LDS (a gather: no atomics), and each endpoint's list comes out in
// increasing rank.  One merge per column then writes the rows in increasing dual order.
//
// Apparent pairs and clearing need no pass over the rows: the leftmost column of triangle row
// t is its latest edge (written with the own entries), that of edge row (a, b), a < b, is
// vertex b; a column leftmost in some row is negative (Lemma L in the dual).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ssq_internal.h"
#include "ssq_kernels.h"
#include "ssq_sort.h"

namespace ssq {

namespace {

constexpr int kMaxWords = kRipsMaxPoints / 64;  // 64-bit words of a vertex bitset
constexpr int kBlock = 256;
constexpr int kWaves = kBlock / kWave;

// Per-group device tables, passed by value.
struct Grp {
  int G;               // clouds in the group
  int64_t NV, NE, NT;  // vertices, edges, triangles of the group
  const int64_t* pb;   // [G+1] first point (vertex) of each cloud
  const int64_t* eb;   // [G+1] first ranked edge
  const int64_t* rb;   // [G+1] rank matrix base (n^2 entries per cloud)
  const int64_t* bb;   // [G+1] Q-bitset word base (E * W per cloud)
  const int64_t* ab;   // [G+1] adjacency-bitset word base (n * W per cloud)
  const int64_t* tb;   // [G+1] first triangle (prefix of the clouds' triangle counts)
};

__device__ __forceinline__ int cloud_of(const int64_t* base, int G, int64_t i) {
  int lo = 0, hi = G;  // largest c with base[c] <= i
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int words_of(int64_t n) { return (int)((n + 63) >> 6); }
__device__ __forceinline__ uint64_t lanes_below() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}
// dual ids of the layout [vertices | edges | triangles]
__device__ __forceinline__ int64_t vdid(const Grp& q, int g, int64_t v_local) {
  return q.pb[g] + (q.pb[g + 1] - q.pb[g] - 1 - v_local);
}
__device__ __forceinline__ int64_t edid(const Grp& q, int g, int64_t k) {
  return q.NV + q.eb[g] + (q.eb[g + 1] - q.eb[g] - 1 - k);
}
__device__ __forceinline__ int64_t tdid(const Grp& q, int g, int64_t rho) {
  return q.NV + q.NE + q.tb[g] + (q.tb[g + 1] - q.tb[g] - 1 - rho);
}

// ---- sizing: adjacency bitsets, edge and triangle counts ---------------------------------
// One wavefront per vertex: A(v) word w, bit x%64 set when x != v is within r; the number of
// neighbours above v (edges (v, x), x > v) and of all neighbours.
__global__ void __launch_bounds__(kBlock) k_rc_adj(const float* P, int pdim, int C, const int64_t* pbase,
                                                   const float* rad, const int64_t* abase, uint64_t* A, int32_t* updeg,
                                                   int32_t* deg) {
  const int lane = lane_id();
  const int64_t nv = pbase[C];
  for (int64_t v = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave; v < nv; v += (int64_t)gridDim.x * kWaves) {
    const int c = cloud_of(pbase, C, v);
    const int64_t b0 = pbase[c], n = pbase[c + 1] - b0, vl = v - b0;
    const int W = words_of(n);
    const float r = rad[c];
    uint64_t* row = A + abase[c] + vl * W;
    int up = 0, all = 0;
    for (int w = 0; w < W; w++) {
      const int64_t x = (int64_t)w * 64 + lane;
      const bool in = x < n && x != vl && rips_dist(P, v, b0 + x, pdim) <= r;
      const uint64_t m = __ballot(in);
      if (lane == 0) row[w] = m;
      all += __popcll(m);
      if (w == (vl >> 6)) up += __popcll(m & ~((2ull << (vl & 63)) - 1));  // bits above vl
      else if (w > (vl >> 6)) up += __popcll(m);
    }
    if (lane == 0) {
      updeg[v] = up;
      deg[v] = all;
    }
  }
}

// One wavefront per vertex v: sum over its edges (v, x), x > v, of |A(v) & A(x)| (each
// triangle is counted once per edge, three times in all).
__global__ void __launch_bounds__(kBlock) k_rc_tri3(int C, const int64_t* pbase, const int64_t* abase,
                                                    const uint64_t* A, int64_t* tri3) {
  const int lane = lane_id();
  const int64_t nv = pbase[C];
  for (int64_t v = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave; v < nv; v += (int64_t)gridDim.x * kWaves) {
    const int c = cloud_of(pbase, C, v);
    const int64_t b0 = pbase[c], n = pbase[c + 1] - b0, vl = v - b0;
    const int W = words_of(n);
    const uint64_t* av = A + abase[c] + vl * W;
    const uint64_t mine = lane < W ? av[lane] : 0ull;
    int64_t t = 0;
    for (int w = (int)(vl >> 6); w < W; w++) {
      uint64_t m = __shfl(mine, w, kWave);
      if (w == (vl >> 6)) m &= ~((2ull << (vl & 63)) - 1);
      while (m) {
        const int bit = __ffsll((unsigned long long)m) - 1;
        m &= m - 1;
        const int64_t x = (int64_t)w * 64 + bit;
        const uint64_t ax = lane < W ? A[abase[c] + x * W + lane] : 0ull;
        t += __popcll(mine & ax);
      }
    }
    for (int d = kWave / 2; d > 0; d >>= 1) t += __shfl_xor(t, d, kWave);
    if (lane == 0) tri3[v] = t;
  }
}

// ---- edges ------------------------------------------------------------------------------
// One wavefront per vertex: its edges (v, x), x > v, in lexicographic order, with sort keys
// cloud << 31 | float bits of the length (lengths are >= 0, so their bits order like them).
__global__ void __launch_bounds__(kBlock) k_rc_efill(Grp q, const float* P, int pdim, const uint64_t* A,
                                                     const int64_t* eoff, int32_t* erow, int32_t* enbr,
                                                     uint64_t* keys, int32_t* vals) {
  const int lane = lane_id();
  for (int64_t v = (int64_t)blockIdx.x * kWaves + threadIdx.x / kWave; v < q.NV; v += (int64_t)gridDim.x * kWaves) {
    const int g = cloud_of(q.pb, q.G, v);
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0, vl = v - b0;
    const int W = words_of(n);
    const uint64_t* av = A + q.ab[g] + vl * W;
    int64_t o = eoff[v];
    for (int w = (int)(vl >> 6); w < W; w++) {
      uint64_t m = av[w];
      if (w == (vl >> 6)) m &= ~((2ull << (vl & 63)) - 1);
      const bool in = (m >> lane) & 1ull;
      if (in) {
        const int64_t x = (int64_t)w * 64 + lane;
        const int64_t e = o + __popcll(m & lanes_below());
        const float d = rips_dist(P, v, b0 + x, pdim);
        erow[e] = (int32_t)v;
        enbr[e] = (int32_t)(b0 + x);
        keys[e] = ((uint64_t)g << 31) | __float_as_uint(d);
        vals[e] = (int32_t)e;
      }
      o += __popcll(m);
    }
  }
}

// Ranked edge p (sorted position): endpoints, length bits, rank matrix, run starts, and the
// dual init of edge rows: the leftmost column of row (a, b), a < b, is vertex b's.
__global__ void k_rc_edges(Grp q, const int32_t* sval, const int32_t* erow, const int32_t* enbr,
                           const uint64_t* skeys, uint32_t* suv, int32_t* ecl, int32_t* R, int32_t* rstart,
                           uint32_t* leftone, uint8_t* negflag) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < q.NE; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t e = sval[p];
    const int64_t i = erow[e], j = enbr[e];
    const int g = cloud_of(q.pb, q.G, i);
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0;
    const int64_t a = i - b0, b = j - b0, k = p - q.eb[g];
    suv[p] = (uint32_t)a | ((uint32_t)b << 16);
    ecl[p] = g;
    R[q.rb[g] + a * n + b] = (int32_t)k;
    R[q.rb[g] + b * n + a] = (int32_t)k;
    rstart[p] = (k == 0 || (uint32_t)skeys[p - 1] != (uint32_t)skeys[p]) ? 1 : 0;
    const int64_t vb = vdid(q, g, b);
    leftone[edid(q, g, k)] = (uint32_t)vb;
    negflag[vb] = 1;
  }
}

// run bounds [rlo, rhi) of every ranked edge from the inclusive count of run starts
__global__ void k_rc_runpos(int64_t ne, const int32_t* rstart, const int32_t* runincl, int32_t* runpos) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x)
    if (rstart[p]) runpos[runincl[p] - 1] = (int32_t)p;
}
__global__ void k_rc_runs(int64_t ne, const int32_t* runincl, const int32_t* runpos, int32_t* rlo, int32_t* rhi) {
  const int32_t nruns = runincl[ne - 1];
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t r = runincl[p] - 1;
    rlo[p] = runpos[r];
    rhi[p] = r + 1 < nruns ? runpos[r + 1] : (int32_t)ne;
  }
}

// Vertex-column lengths in dual order (vertex column vdid(v) holds v's incident edges), and
// each vertex's checkpoint words (one bitset per 64 neighbours).
__global__ void k_rc_vlens(Grp q, const int32_t* deg, int64_t* lens, int64_t* cpw) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < q.NV; c += (int64_t)gridDim.x * blockDim.x) {
    const int g = cloud_of(q.pb, q.G, c);
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0;
    lens[c] = deg[b0 + (n - 1 - (c - b0))];
    cpw[c] = (int64_t)((deg[c] + 63) / 64) * words_of(n);
  }
}

// ---- per vertex: incident edges by rank, checkpoints, the vertex's dual column ---------------
// One workgroup per vertex v.  The incident edges sorted by rank give: inc (rank, neighbour),
// epos (the edge's position in each endpoint's list = how many of that endpoint's edges are
// older), checkpoint bitsets CP[j] = the first 64(j+1) neighbours, and the vertex's coboundary
// column (its edges, newest first = increasing dual id).
__global__ void __launch_bounds__(kBlock) k_rc_vertex(Grp q, const uint64_t* A, const int32_t* R, const int64_t* aoff,
                                                      const int64_t* cpoff, const int64_t* colptr, int32_t* inc_r,
                                                      uint16_t* inc_n, int32_t* epos, uint64_t* CP, int32_t* pool) {
  __shared__ uint64_t ent[kRipsMaxPoints];
  __shared__ uint64_t bits[kMaxWords];
  __shared__ int32_t fill;
  const int t = threadIdx.x;
  for (int64_t v = blockIdx.x; v < q.NV; v += gridDim.x) {
    const int g = cloud_of(q.pb, q.G, v);
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0, vl = v - b0;
    const int W = words_of(n);
    const int64_t ao = aoff[v];
    const int deg = (int)(aoff[v + 1] - ao);
    const int32_t* row = R + q.rb[g] + vl * n;
    const uint64_t* av = A + q.ab[g] + vl * W;
    if (t == 0) fill = 0;
    __syncthreads();
    // neighbours (rank << 12 | id); a wavefront per 64-word stretch, positions by ballot
    for (int64_t x0 = 0; x0 < n; x0 += kBlock) {
      const int64_t x = x0 + t;
      const bool in = x < n && ((av[x >> 6] >> (x & 63)) & 1ull);
      const uint64_t m = __ballot(in);
      int32_t base = 0;
      if (lane_id() == 0 && m) base = atomicAdd(&fill, (int32_t)__popcll(m));
      base = __shfl(base, 0, kWave);
      if (in) ent[base + __popcll(m & lanes_below())] = ((uint64_t)(uint32_t)row[x] << 12) | (uint64_t)x;
    }
    __syncthreads();
    int P2 = 1;
    while (P2 < deg) P2 <<= 1;
    for (int i = deg + t; i < P2; i += kBlock) ent[i] = ~0ull;
    __syncthreads();
    for (int kk = 2; kk <= P2; kk <<= 1)  // bitonic sort, ascending
      for (int jj = kk >> 1; jj > 0; jj >>= 1) {
        for (int i = t; i < P2; i += kBlock) {
          const int ixj = i ^ jj;
          if (ixj > i) {
            const uint64_t x = ent[i], y = ent[ixj];
            if ((x > y) == ((i & kk) == 0)) {
              ent[i] = y;
              ent[ixj] = x;
            }
          }
        }
        __syncthreads();
      }
    const int64_t Eg = q.eb[g + 1] - q.eb[g];
    const int64_t col = colptr[vdid(q, g, vl)];
    for (int i = t; i < deg; i += kBlock) {
      const uint64_t x = ent[i];
      const int32_t k = (int32_t)(x >> 12);
      const int32_t nb = (int32_t)(x & 0xfff);
      inc_r[ao + i] = k;
      inc_n[ao + i] = (uint16_t)nb;
      const int64_t p = q.eb[g] + k;
      epos[2 * p + (vl > nb ? 1 : 0)] = i;
      pool[col + (deg - 1 - i)] = (int32_t)(q.NV + q.eb[g] + (Eg - 1 - k));
    }
    if (t < W) bits[t] = 0;
    __syncthreads();
    const int nchunk = (deg + 63) >> 6;
    for (int j = 0; j < nchunk; j++) {
      if (t < 64 && j * 64 + t < deg) {
        const int x = (int)(ent[j * 64 + t] & 0xfff);
        atomicOr((unsigned long long*)&bits[x >> 6], 1ull << (x & 63));
      }
      __syncthreads();
      if (t < W) CP[cpoff[v] + (int64_t)j * W + t] = bits[t];
      __syncthreads();
    }
  }
}

// ---- Q bitsets ----------------------------------------------------------------------------
// One wavefront per ranked edge m = (a, b) of rank k; lane w holds word w.  Q(m) = (older
// neighbours of a) & (older neighbours of b): a checkpoint plus the partial chunk of each
// endpoint's rank-sorted list.  Also |Q| (own triangles), the word prefix counts and the number
// of common neighbours (the coboundary's length).
__global__ void __launch_bounds__(kBlock) k_rc_qbits(Grp q, const uint32_t* suv, const int32_t* ecl,
                                                     const int32_t* epos, const int64_t* aoff, const int64_t* cpoff,
                                                     const uint16_t* inc_n, const uint64_t* CP, const uint64_t* A,
                                                     uint64_t* B, uint16_t* Bp, int64_t* cnt, int32_t* cof,
                                                     int64_t* lens, uint8_t* negflag) {
  __shared__ uint64_t wa[kWaves][kMaxWords], wb[kWaves][kMaxWords];
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  for (int64_t p = (int64_t)blockIdx.x * kWaves + wv; p < q.NE; p += (int64_t)gridDim.x * kWaves) {
    const int g = ecl[p];
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0;
    const int W = words_of(n);
    const int64_t a = suv[p] & 0xffff, b = suv[p] >> 16, k = p - q.eb[g];
    const int64_t va = b0 + a, vb = b0 + b;
    const int pa = epos[2 * p], pbp = epos[2 * p + 1];
    uint64_t fa = 0, fb = 0;
    if (lane < W) {
      wa[wv][lane] = (pa >> 6) ? CP[cpoff[va] + (int64_t)((pa >> 6) - 1) * W + lane] : 0ull;
      wb[wv][lane] = (pbp >> 6) ? CP[cpoff[vb] + (int64_t)((pbp >> 6) - 1) * W + lane] : 0ull;
      fa = A[q.ab[g] + a * W + lane];
      fb = A[q.ab[g] + b * W + lane];
    }
    wave_sync();
    if (lane < (pa & 63)) {
      const int x = inc_n[aoff[va] + (pa & ~63) + lane];
      atomicOr((unsigned long long*)&wa[wv][x >> 6], 1ull << (x & 63));
    }
    if (lane < (pbp & 63)) {
      const int x = inc_n[aoff[vb] + (pbp & ~63) + lane];
      atomicOr((unsigned long long*)&wb[wv][x >> 6], 1ull << (x & 63));
    }
    wave_sync();
    const uint64_t qw = lane < W ? (wa[wv][lane] & wb[wv][lane]) : 0ull;
    const int c = __popcll(qw);
    const int incl = wave_incl_scan(c);
    int common = __popcll(fa & fb);
    for (int d = kWave / 2; d > 0; d >>= 1) common += __shfl_xor(common, d, kWave);
    const int total = __shfl(incl, kWave - 1, kWave);
    if (lane < W) {
      B[q.bb[g] + k * W + lane] = qw;
      Bp[q.bb[g] + k * W + lane] = (uint16_t)(incl - c);
    }
    if (lane == 0) {
      cnt[p] = total;
      cof[p] = common;
      const int64_t ec = edid(q, g, k);
      lens[ec] = common;
      if (total > 0) negflag[ec] = 1;
    }
    wave_sync();
  }
}

// ---- runs of equal-length edges: lexicographic ranks of their triangles ----------------------
__global__ void k_rc_tiecnt(int64_t ne, const int32_t* rlo, const int32_t* rhi, const int64_t* cnt, int64_t* tcnt,
                            int32_t* tiestart) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x) {
    const bool tie = rhi[p] - rlo[p] > 1;
    tcnt[p] = tie ? cnt[p] : 0;
    tiestart[p] = (tie && rlo[p] == p) ? 1 : 0;
  }
}

// One wavefront per edge of a tie run: its own triangles with key (run index, sorted vertex
// triple) and value the natural index TB[m] + position in Q(m).
__global__ void __launch_bounds__(kBlock) k_rc_tieenum(Grp q, const uint32_t* suv, const int32_t* ecl,
                                                       const int32_t* rlo, const int32_t* rhi, const int64_t* tidx_excl,
                                                       const int64_t* toff, const int64_t* TB, const uint64_t* B,
                                                       const uint16_t* Bp, uint64_t* keys, int32_t* vals) {
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  for (int64_t p = (int64_t)blockIdx.x * kWaves + wv; p < q.NE; p += (int64_t)gridDim.x * kWaves) {
    if (rhi[p] - rlo[p] <= 1) continue;
    const int g = ecl[p];
    const int64_t n = q.pb[g + 1] - q.pb[g];
    const int W = words_of(n);
    const uint32_t a = suv[p] & 0xffff, b = suv[p] >> 16;
    const int64_t k = p - q.eb[g];
    const uint64_t run = (uint64_t)tidx_excl[rlo[p]];
    if (lane >= W) continue;
    uint64_t m = B[q.bb[g] + k * W + lane];
    int i = Bp[q.bb[g] + k * W + lane];
    while (m) {
      const uint32_t z = (uint32_t)(lane * 64 + __ffsll((unsigned long long)m) - 1);
      m &= m - 1;
      uint32_t x0 = a, x1 = b, x2 = z;  // a < b; place z
      if (z < a) {
        x0 = z;
        x1 = a;
        x2 = b;
      } else if (z < b) {
        x1 = z;
        x2 = b;
      }
      const int64_t j = toff[p] + i;
      keys[j] = (run << 36) | ((uint64_t)x0 << 24) | ((uint64_t)x1 << 12) | (uint64_t)x2;
      vals[j] = (int32_t)(TB[p] + i);
      i++;
    }
  }
}

// Run r's triangles are at sorted positions [toff[first edge], ...): remap the natural index
// to the triangle's cloud-local rank.
__global__ void k_rc_tieruns(int64_t ne, const int32_t* tiestart, const int64_t* tidx_excl, const int64_t* toff,
                             const int64_t* TB, const int32_t* ecl, const int64_t* tbase, int64_t* run_toff,
                             int64_t* run_rank0) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < ne; p += (int64_t)gridDim.x * blockDim.x)
    if (tiestart[p]) {
      run_toff[tidx_excl[p]] = toff[p];
      run_rank0[tidx_excl[p]] = TB[p] - tbase[ecl[p]];
    }
}
__global__ void k_rc_tieapply(int64_t nt, const uint64_t* skeys, const int32_t* svals, const int64_t* run_toff,
                              const int64_t* run_rank0, int32_t* remap) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nt; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = (int64_t)(skeys[j] >> 36);
    remap[svals[j]] = (int32_t)(run_rank0[r] + (j - run_toff[r]));
  }
}

// ---- the two endpoint lists of every coboundary column ------------------------------------
// One workgroup per vertex v.  For each word block w, the block stages word w of Q(m), its
// prefix count and TB[m] for every edge m = (v, y) of v (rank order).  Column (v, z), z in the
// block, takes the triangles {v, y, z} whose latest edge is (v, y): the y later than (v, z)
// with z in Q((v, y)), rank TB + prefix + popcount below z.  Scanning y by rank yields them in
// increasing rank.  v the lower endpoint: the list fills the column's foreign region from its
// start; the higher: from its end, backwards.
__global__ void __launch_bounds__(kBlock) k_rc_side(Grp q, const int64_t* aoff, const int32_t* inc_r,
                                                    const uint16_t* inc_n, const uint64_t* B, const uint16_t* Bp,
                                                    const int64_t* TB, const int32_t* rlo, const int32_t* rhi,
                                                    const int32_t* remap, const int64_t* colptr, const int64_t* cnt,
                                                    const int32_t* cof, int32_t* stg, int32_t* na, int32_t* nb) {
  __shared__ uint64_t sw[kRipsMaxPoints];
  __shared__ int32_t stb[kRipsMaxPoints];  // cloud-local TB[m], sign bit: m's run has several edges
  __shared__ uint16_t sp[kRipsMaxPoints];
  __shared__ int16_t zpos[kRipsMaxPoints];
  const int t = threadIdx.x, lane = lane_id(), wv = t / kWave;
  const int64_t stg0 = 2 * q.NE;  // the edge columns' rows start after the vertex columns'
  for (int64_t v = blockIdx.x; v < q.NV; v += gridDim.x) {
    const int g = cloud_of(q.pb, q.G, v);
    const int64_t b0 = q.pb[g], n = q.pb[g + 1] - b0, vl = v - b0;
    const int W = words_of(n);
    const int64_t ao = aoff[v], e0 = q.eb[g], Eg = q.eb[g + 1] - e0, tbc = q.tb[g];
    const int deg = (int)(aoff[v + 1] - ao);
    for (int x = t; x < n; x += kBlock) zpos[x] = -1;
    __syncthreads();
    for (int i = t; i < deg; i += kBlock) zpos[inc_n[ao + i]] = (int16_t)i;
    for (int w = 0; w < W; w++) {
      __syncthreads();
      for (int i = t; i < deg; i += kBlock) {
        const int64_t k = inc_r[ao + i], p = e0 + k;
        sw[i] = B[q.bb[g] + k * W + w];
        sp[i] = Bp[q.bb[g] + k * W + w];
        const uint32_t tie = rhi[p] - rlo[p] > 1 ? 0x80000000u : 0u;
        stb[i] = (int32_t)((uint32_t)(TB[p] - tbc) | tie);
      }
      __syncthreads();
      for (int zz = wv; zz < 64; zz += kWaves) {
        const int z = w * 64 + zz;
        if (z >= n) break;
        const int iz = zpos[z];
        if (iz < 0) continue;
        const int64_t kz = inc_r[ao + iz], pe = e0 + kz;
        const int64_t cb = colptr[q.NV + e0 + (Eg - 1 - kz)] - stg0;
        const int64_t F = cof[pe] - cnt[pe];
        const bool lower = vl < z;
        const uint64_t below = (1ull << zz) - 1;
        int64_t written = 0;
        for (int i0 = iz + 1; i0 < deg; i0 += kWave) {
          const int i = i0 + lane;
          bool hit = false;
          int32_t rho = 0;
          if (i < deg) {
            const uint64_t x = sw[i];
            if ((x >> zz) & 1ull) {
              hit = true;
              const uint32_t s = (uint32_t)stb[i];
              rho = (int32_t)(s & 0x7fffffffu) + sp[i] + __popcll(x & below);
              if (s & 0x80000000u) rho = remap[tbc + rho];
            }
          }
          const uint64_t m = __ballot(hit);
          if (hit) {
            const int64_t pos = written + __popcll(m & lanes_below());
            if (pos < F) stg[cb + (lower ? pos : F - 1 - pos)] = rho;  // more than F: reported by the merge
          }
          written += __popcll(m);
        }
        if (lane == 0) (lower ? na : nb)[pe] = (int32_t)written;
      }
    }
    __syncthreads();
  }
}

// ---- merge: the coboundary column in increasing dual id (decreasing rank) ------------------
// One wavefront per edge column, three sorted lists: the lower endpoint's (ascending), the
// higher endpoint's (stored descending) and the own triangles (ascending: their natural index,
// through the run's lexicographic remap when the edge's length is shared).  Within a run of
// equal-length edges the lists stay sorted (the triangles of a fixed vertex pair are
// lexicographically ordered by the third vertex, and so are a vertex's edges within a run);
// only the own list interleaves with the foreign entries of its run.  Each entry is placed by
// counting the larger entries of the other two lists (binary searches).  Each own triangle's
// leftmost column is this one.
__device__ __forceinline__ int32_t own_rank(const int32_t* remap, bool tie, int64_t base, int64_t own0, int i) {
  return tie ? remap[base + own0 + i] : (int32_t)(own0 + i);
}
// number of entries > x in an ascending list / a descending list / the own list
__device__ __forceinline__ int count_gt_asc(const int32_t* v, int n, int32_t x) {
  int lo = 0, hi = n;  // first index with v > x
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return n - lo;
}
__device__ __forceinline__ int count_gt_desc(const int32_t* v, int n, int32_t x) {
  int lo = 0, hi = n;  // first index with v < x (v > x before it)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] > x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int count_gt_own(const int32_t* remap, bool tie, int64_t base, int64_t own0, int n,
                                            int32_t x) {
  if (!tie) {  // own0 .. own0 + n - 1
    const int64_t c = own0 + n - 1 - x;
    return (int)(c < 0 ? 0 : (c > n ? n : c));
  }
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (remap[base + own0 + mid] <= x) lo = mid + 1;
    else hi = mid;
  }
  return n - lo;
}

__global__ void __launch_bounds__(kBlock) k_rc_merge(Grp q, const int32_t* ecl, const int32_t* rlo, const int32_t* rhi,
                                                     const int64_t* TB, const int32_t* remap, const int64_t* colptr,
                                                     const int64_t* cnt, const int32_t* cof, const int32_t* na,
                                                     const int32_t* nb, const int32_t* stg, int32_t* pool,
                                                     uint32_t* leftone, uint32_t* err) {
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  const int64_t stg0 = 2 * q.NE;
  for (int64_t p = (int64_t)blockIdx.x * kWaves + wv; p < q.NE; p += (int64_t)gridDim.x * kWaves) {
    const int g = ecl[p];
    const int64_t e0 = q.eb[g], k = p - e0, tbc = q.tb[g];
    const int64_t col = edid(q, g, k);
    const int64_t c0 = colptr[col];
    const int32_t len = cof[p], own = (int32_t)cnt[p], F = len - own;
    const int32_t an = na[p], bn = nb[p];
    if (an + bn != F || an < 0 || bn < 0 || own < 0) {
      if (lane == 0) atomicOr(err, 1u);
      continue;
    }
    const bool tie = own > 0 && rhi[p] - rlo[p] > 1;
    const int32_t* Aa = stg + (c0 - stg0);  // ascending, an entries
    const int32_t* Bd = Aa + (F - bn);      // descending, bn entries
    int32_t* out = pool + c0;
    const int64_t own0 = TB[p] - tbc;
    for (int i = lane; i < an; i += kWave) {
      const int32_t x = Aa[i];
      const int pos = (an - 1 - i) + count_gt_desc(Bd, bn, x) + count_gt_own(remap, tie, tbc, own0, own, x);
      out[pos] = (int32_t)tdid(q, g, x);
    }
    for (int j = lane; j < bn; j += kWave) {
      const int32_t y = Bd[j];
      const int pos = j + count_gt_asc(Aa, an, y) + count_gt_own(remap, tie, tbc, own0, own, y);
      out[pos] = (int32_t)tdid(q, g, y);
    }
    for (int i = lane; i < own; i += kWave) {
      const int32_t z = own_rank(remap, tie, tbc, own0, i);
      const int pos = (own - 1 - i) + count_gt_asc(Aa, an, z) + count_gt_desc(Bd, bn, z);
      const int64_t d = tdid(q, g, z);
      out[pos] = (int32_t)d;
      leftone[d] = (uint32_t)col;
    }
  }
}

// Every column's rows strictly increasing and below n (the engine indexes its tables with
// them): a check of the built matrix before the reduction reads it.  One wavefront per column.
__global__ void __launch_bounds__(kBlock) k_rc_check(int64_t nc, int64_t n, const int64_t* colptr, const int32_t* pool,
                                                     uint32_t* err) {
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  for (int64_t c = (int64_t)blockIdx.x * kWaves + wv; c < nc; c += (int64_t)gridDim.x * kWaves) {
    const int64_t o = colptr[c], len = colptr[c + 1] - o;
    bool bad = false;
    for (int64_t i = lane; i < len; i += kWave) {
      const int32_t x = pool[o + i];
      bad |= x < 0 || x >= n || (i > 0 && pool[o + i - 1] >= x);
    }
    if (__ballot(bad) && lane == 0) atomicOr(err, 2u);
  }
}

// ---- pairs --------------------------------------------------------------------------------
// Columns in primal order (cloud-major vertices, then cloud-major edges by rank): 1 where the
// column ends reduced and owns its low row.
__device__ __forceinline__ int64_t enum_col(const Grp& q, const int32_t* ecl, int64_t i, int* g_out) {
  if (i < q.NV) {
    const int g = cloud_of(q.pb, q.G, i);
    *g_out = g;
    return vdid(q, g, i - q.pb[g]);
  }
  const int64_t p = i - q.NV;
  const int g = ecl[p];
  *g_out = g;
  return edid(q, g, p - q.eb[g]);
}
__global__ void k_rc_pairflag(Grp q, const int32_t* ecl, const int32_t* low, const uint32_t* piv, int32_t* flag) {
  const int64_t nc = q.NV + q.NE;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
    int g;
    const int64_t c = enum_col(q, ecl, i, &g);
    const int32_t l = low[c];
    flag[i] = (l >= 0 && piv[l] == (uint32_t)c) ? 1 : 0;
  }
}
// per cloud: pairs with a vertex birth, with an edge birth (from the scan of the flags)
__global__ void k_rc_paircount(Grp q, const int64_t* pos, int64_t* hv, int64_t* he) {
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < q.G; g += gridDim.x * blockDim.x) {
    hv[g] = pos[q.pb[g + 1]] - pos[q.pb[g]];
    he[g] = pos[q.NV + q.eb[g + 1]] - pos[q.NV + q.eb[g]];
  }
}
// primal cell of ranked edge p: n + rank + triangles of smaller filtration
__device__ __forceinline__ int64_t edge_cell(const Grp& q, int g, int64_t p, const int32_t* rlo, const int64_t* TB) {
  return (q.pb[g + 1] - q.pb[g]) + (p - q.eb[g]) + (TB[rlo[p]] - q.tb[g]);
}
__global__ void k_rc_pairemit(Grp q, const int32_t* ecl, const int32_t* low, const int32_t* flag, const int64_t* pos,
                              const int64_t* hv, const int64_t* ob, const int32_t* rlo, const int32_t* rhi,
                              const int64_t* TB, int64_t* out) {
  const int64_t nc = q.NV + q.NE;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[i]) continue;
    int g;
    const int64_t c = enum_col(q, ecl, i, &g);
    const int64_t r = low[c];
    const int64_t n = q.pb[g + 1] - q.pb[g], e0 = q.eb[g], Eg = q.eb[g + 1] - e0;
    int64_t birth, death, at;
    if (i < q.NV) {
      birth = i - q.pb[g];
      at = ob[g] + (pos[i] - pos[q.pb[g]]);
      const int64_t k = Eg - 1 - (r - q.NV - e0);  // the low row is an edge of this cloud
      death = edge_cell(q, g, e0 + k, rlo, TB);
    } else {
      const int64_t p = i - q.NV;
      birth = edge_cell(q, g, p, rlo, TB);
      at = ob[g] + hv[g] + (pos[i] - pos[q.NV + e0]);
      const int64_t tbc = q.tb[g], Tg = q.tb[g + 1] - tbc;
      const int64_t rho = Tg - 1 - (r - q.NV - q.NE - tbc);
      int64_t lo = e0, hi = q.eb[g + 1];  // last edge with local TB <= rho
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (TB[mid] - tbc <= rho) lo = mid;
        else hi = mid;
      }
      death = n + (rhi[lo] - e0) + rho;
    }
    out[2 * at] = birth;
    out[2 * at + 1] = death;
  }
}

// A group's inputs gathered from the sizing pass's arrays (clouds in any order): points,
// degrees and adjacency bitsets of group cloud g come from its vertex base srcv[g] and bitset
// base srca[g].
__global__ void k_rc_gather(Grp q, int pdim, const int64_t* srcv, const int64_t* srca, const float* Pall,
                            const uint64_t* Aall, const int32_t* degall, const int32_t* upall, float* P, uint64_t* A,
                            int32_t* deg, int32_t* up) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < q.NV; v += (int64_t)gridDim.x * blockDim.x) {
    const int g = cloud_of(q.pb, q.G, v);
    const int64_t vl = v - q.pb[g], sv = srcv[g] + vl;
    const int W = words_of(q.pb[g + 1] - q.pb[g]);
    for (int k = 0; k < pdim; k++) P[v * pdim + k] = Pall[sv * pdim + k];
    deg[v] = degall[sv];
    up[v] = upall[sv];
    const uint64_t* a = Aall + srca[g] + vl * W;
    uint64_t* o = A + q.ab[g] + vl * W;
    for (int w = 0; w < W; w++) o[w] = a[w];
  }
}

// Pairs from per-cloud slots to their packed place: cloud c's cnt[c] pairs move from slot[c]
// to off[c].  One workgroup per cloud.
__global__ void k_rc_pack(int C, const int64_t* slot, const int64_t* off, const int64_t* stage, int64_t* out) {
  for (int c = blockIdx.x; c < C; c += gridDim.x) {
    const int64_t n = 2 * (off[c + 1] - off[c]);
    const int64_t* src = stage + 2 * slot[c];
    int64_t* dst = out + 2 * off[c];
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
  }
}

__global__ void k_rc_set64(int64_t* p, int64_t v) { *p = v; }

__global__ void k_rc_dims(int64_t nv, int64_t nc, int32_t* dims) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    dims[c] = c < nv ? 2 : 1;  // dual dimension: vertices 2, edges 1
}

__global__ void k_rc_fill32(int64_t n, uint32_t* p, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace

// ====================================================================================
// host side
// ====================================================================================
struct RipsCo {
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  // sizing (all clouds) and per-group buffers; grown on demand and kept
  Buf pb, rad, ab, A, updeg, deg, tri3, temp;
  Buf gtab, eoff, erow, enbr, keys, keys2, vals, vals2, suv, ecl, R, rstart, runincl, runpos, rlo, rhi, lens,
      colptr, aoff, cpoff, cpw, inc_r, inc_n, epos, CP, B, Bp, cnt, cof, TB, tcnt, toff, tiestart, tidx, run_toff,
      run_rank0, remap, tkeys, tkeys2, tvals, tvals2, na, nb, dims, flag, pos, hv, he, err, pairs, gpts, gA, gdeg,
      gup, stage, slots;
  int64_t* host = nullptr;  // pinned scratch for read-backs
  int64_t budget = 0;       // device bytes the streams' groups may use together
  int budget_streams = 0;   // the streams it was split over (a group gets budget / budget_streams)
  bool trace = std::getenv("SSQ_TRACE_GROUPS") != nullptr;  // per-group sizes and times on stderr
  std::vector<int64_t> h_pb, h_ab;
  std::vector<float> h_rad;
  std::vector<int64_t> h_edges, h_tris;
  ~RipsCo() {
    for (Buf* b : {&pb, &rad, &ab, &A, &updeg, &deg, &tri3, &temp, &gtab, &eoff, &erow, &enbr, &keys, &keys2, &vals,
                   &vals2, &suv, &ecl, &R, &rstart, &runincl, &runpos, &rlo, &rhi, &lens, &colptr, &aoff, &cpoff, &cpw,
                   &inc_r, &inc_n, &epos, &CP, &B, &Bp, &cnt, &cof, &TB, &tcnt, &toff, &tiestart, &tidx, &run_toff,
                   &run_rank0, &remap, &tkeys, &tkeys2, &tvals, &tvals2, &na, &nb, &dims, &flag, &pos, &hv,
                   &he, &err, &pairs, &gpts, &gA, &gdeg, &gup, &stage, &slots})
      if (b->p) (void)hipFree(b->p);
    if (host) (void)hipHostFree(host);
  }
};

RipsCo* ripsco_create() {
  RipsCo* b = new (std::nothrow) RipsCo();
  if (b && hipHostMalloc((void**)&b->host, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return nullptr;
  }
  return b;
}
void ripsco_destroy(RipsCo* b) { delete b; }
const int64_t* ripsco_triangles(const RipsCo* b) { return b->h_tris.data(); }
const int64_t* ripsco_edges(const RipsCo* b) { return b->h_edges.data(); }

namespace {

template <typename T>
bool rc_grow(RipsCo::Buf& b, int64_t count, T** out) {
  const size_t need = (size_t)std::max<int64_t>(count, 1) * sizeof(T) + 256;
  if (b.bytes < need) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    const size_t want = need + need / 4;  // a free synchronizes the device: grow with headroom
    if (hipMalloc(&b.p, want) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    b.bytes = want;
  }
  *out = (T*)b.p;
  return true;
}

int grid_of(int64_t n, int block = kBlock) {
  return (int)std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, 16384));
}

#define RCCHK(x)                                    \
  do {                                              \
    if ((x) != hipSuccess) {                        \
      (void)hipGetLastError();                      \
      if (err) *err = std::string("HIP error: ") + #x; \
      return SSQ_EDEVICE;                           \
    }                                               \
  } while (0)
#define RCGROW(buf, T, n, ptr)                                \
  do {                                                        \
    if (!rc_grow<T>(B->buf, (int64_t)(n), &(ptr))) {          \
      if (err) *err = "device allocation failed (" #buf ")"; \
      return SSQ_ENOMEM;                                      \
    }                                                         \
  } while (0)

// exclusive scan (int64) of n values; out[n] = total
template <typename In>
int scan64(RipsCo* B, hipStream_t s, const In* in, int64_t n, int64_t* out, std::string* err) {
  if (n == 0) {
    hipLaunchKernelGGL(k_rc_set64, dim3(1), dim3(1), 0, s, out, (int64_t)0);
    return SSQ_OK;
  }
  uint8_t* tmp;
  RCGROW(temp, uint8_t, ssq_sort::scan_bytes(n), tmp);
  RCCHK((ssq_sort::scan<In, int64_t, false>(in, n, out, tmp, s)));  // out[n] = the total
  return SSQ_OK;
}

}  // namespace

// Sizing pass over all clouds: adjacency bitsets (kept for the reduction), edges, triangles.
static int rc_size(RipsCo* B, hipStream_t s, int C, const float* d_pts, const int64_t* pbase, int pdim,
                   const float* radius, std::string* err) {
  B->h_pb.assign(pbase, pbase + C + 1);
  B->h_rad.assign(radius, radius + C);
  B->h_ab.assign((size_t)C + 1, 0);
  for (int c = 0; c < C; c++) {
    const int64_t n = pbase[c + 1] - pbase[c];
    if (n < 0 || n > kRipsMaxPoints || !(radius[c] >= 0.f)) {
      if (err) *err = "cloud size or threshold out of range for the batched Rips path";
      return SSQ_EINVAL;
    }
    B->h_ab[c + 1] = B->h_ab[c] + n * ((n + 63) / 64);
  }
  const int64_t nv = pbase[C];
  int64_t *d_pb, *d_ab, *d_tri3;
  float* d_rad;
  uint64_t* A;
  int32_t *updeg, *deg;
  RCGROW(pb, int64_t, C + 1, d_pb);
  RCGROW(ab, int64_t, C + 1, d_ab);
  RCGROW(rad, float, C, d_rad);
  RCGROW(A, uint64_t, B->h_ab[C], A);
  RCGROW(updeg, int32_t, nv, updeg);
  RCGROW(deg, int32_t, nv, deg);
  RCGROW(tri3, int64_t, nv, d_tri3);
  RCCHK(hipMemcpyAsync(d_pb, pbase, sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, s));
  RCCHK(hipMemcpyAsync(d_ab, B->h_ab.data(), sizeof(int64_t) * (C + 1), hipMemcpyHostToDevice, s));
  RCCHK(hipMemcpyAsync(d_rad, radius, sizeof(float) * C, hipMemcpyHostToDevice, s));
  if (nv > 0) {
    const int g = grid_of(nv, kWaves);
    hipLaunchKernelGGL(k_rc_adj, dim3(g), dim3(kBlock), 0, s, d_pts, pdim, C, (const int64_t*)d_pb,
                       (const float*)d_rad, (const int64_t*)d_ab, A, updeg, deg);
    hipLaunchKernelGGL(k_rc_tri3, dim3(g), dim3(kBlock), 0, s, C, (const int64_t*)d_pb, (const int64_t*)d_ab,
                       (const uint64_t*)A, d_tri3);
    RCCHK(hipGetLastError());
  }
  std::vector<int32_t> hup((size_t)nv);
  std::vector<int64_t> ht((size_t)nv);
  if (nv > 0) {
    RCCHK(hipMemcpyAsync(hup.data(), updeg, sizeof(int32_t) * nv, hipMemcpyDeviceToHost, s));
    RCCHK(hipMemcpyAsync(ht.data(), d_tri3, sizeof(int64_t) * nv, hipMemcpyDeviceToHost, s));
  }
  RCCHK(hipStreamSynchronize(s));
  B->h_edges.assign((size_t)C, 0);
  B->h_tris.assign((size_t)C, 0);
  for (int c = 0; c < C; c++) {
    int64_t e = 0, t3 = 0;
    for (int64_t v = pbase[c]; v < pbase[c + 1]; v++) {
      e += hup[(size_t)v];
      t3 += ht[(size_t)v];
    }
    B->h_edges[(size_t)c] = e;
    B->h_tris[(size_t)c] = t3 / 3;
  }
  return SSQ_OK;
}

int ripsco_count(RipsCo* B, hipStream_t s, int count, const float* d_pts, const int64_t* pbase, int pdim,
                 const float* radius, int64_t* edges, int64_t* triangles) {
  if (!B || count < 1 || !pbase || !radius || pdim < 1 || (pbase[count] > 0 && !d_pts)) return SSQ_EINVAL;
  std::string e;
  std::string* err = &e;
  const int rc = rc_size(B, s, count, d_pts, pbase, pdim, radius, err);
  if (rc) return rc;
  for (int c = 0; c < count; c++) {
    if (edges) edges[c] = B->h_edges[(size_t)c];
    if (triangles) triangles[c] = B->h_tris[(size_t)c];
  }
  return SSQ_OK;
}

// Cells of one group: a bound well inside the engine's int32 cell ids, and small enough that
// the per-row tables of two consecutive groups' sizes fit beside each other as buffers grow.
constexpr int64_t kGroupCells = (int64_t)1 << 30;
// device bytes each engine keeps for its huge-column bit sets (k_add_huge regions)
constexpr int64_t kHugeRegionBytes = (int64_t)256 << 24;

// Device bytes of one group, by the buffers below (the engine's included), with the slack of
// their geometric growth (1.5x for the engine's buffers, 1.25x for the builder's).
static double group_bytes(int64_t nv, int64_t ne, int64_t nt, int64_t n2, int64_t ew, int64_t tie_guess) {
  const double rows = 2.0 * ne + 3.0 * nt;
  const double engine = 4.0 * rows * 2.0                // pool: rows, then the arena (holds the staging)
                        + 42.0 * (double)(nv + ne + nt) // per-row tables: piv, leftone, flags, pivmeta
                        + 40.0 * (double)(nv + ne);     // per-column state and lists
  const double builder = 4.0 * (double)nt               // tie remap
                         + 4.0 * (double)n2             // rank matrices
                         + 10.0 * (double)ew            // Q bitsets and prefix counts
                         + 120.0 * (double)ne + 80.0 * (double)nv  // edge keys, runs, lists, lengths
                         + 32.0 * (double)tie_guess;
  return 1.5 * engine + 1.25 * builder + 64.0 * 1048576.0;
}

// One group: clouds[0 .. G) (indices into the sizing pass S), built and reduced on stream s;
// each cloud's pairs go to stage at its slot, their count to counts[cloud].
static int rc_group(RipsCo* B, const RipsCo* S, PrebuiltTarget* eng, hipStream_t s, const int* clouds, int G,
                    const float* d_pts, int pdim, const ssq_params& prm, int64_t* stage, const int64_t* slot,
                    int64_t* counts, int64_t* arena_hint, ssq_stats* acc, std::string* err) {
  auto tb0 = std::chrono::steady_clock::now();
  std::vector<int64_t> pb(G + 1), eb(G + 1), rb(G + 1), bb(G + 1), ab(G + 1), tb(G + 1);
  std::vector<int64_t> srcv(G), srca(G), gslot(G);
  pb[0] = eb[0] = rb[0] = bb[0] = ab[0] = tb[0] = 0;
  for (int g = 0; g < G; g++) {
    const int c = clouds[g];
    const int64_t n = S->h_pb[c + 1] - S->h_pb[c], W = (n + 63) / 64, E = S->h_edges[(size_t)c];
    pb[g + 1] = pb[g] + n;
    ab[g + 1] = ab[g] + n * W;
    eb[g + 1] = eb[g] + E;
    rb[g + 1] = rb[g] + n * n;
    bb[g + 1] = bb[g] + E * W;
    tb[g + 1] = tb[g] + S->h_tris[(size_t)c];
    srcv[g] = S->h_pb[c];
    srca[g] = S->h_ab[c];
    gslot[g] = slot[c];
  }
  const int64_t NV = pb[G], NE = eb[G], NT = tb[G], NC = NV + NE, N = NC + NT;
  const int64_t NNZ = 2 * NE + 3 * NT;
  if (NV == 0) {  // empty clouds only
    for (int g = 0; g < G; g++) counts[clouds[g]] = 0;
    return SSQ_OK;
  }
  if (N >= 0x7fffffffLL) {
    if (err) *err = "group exceeds 2^31 cells";
    return SSQ_EINVAL;
  }
  // group tables on the device: pb | eb | rb | bb | ab | tb | srcv | srca | slot
  int64_t* gt;
  RCGROW(gtab, int64_t, 9 * (G + 1), gt);
  {
    std::vector<int64_t> h(9 * (size_t)(G + 1), 0);
    for (int g = 0; g <= G; g++) {
      h[g] = pb[g];
      h[(G + 1) + g] = eb[g];
      h[2 * (G + 1) + g] = rb[g];
      h[3 * (G + 1) + g] = bb[g];
      h[4 * (G + 1) + g] = ab[g];
      h[5 * (G + 1) + g] = tb[g];
      if (g < G) {
        h[6 * (G + 1) + g] = srcv[g];
        h[7 * (G + 1) + g] = srca[g];
        h[8 * (G + 1) + g] = gslot[g];
      }
    }
    RCCHK(hipMemcpyAsync(gt, h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice, s));
  }
  Grp q;
  q.G = G;
  q.NV = NV;
  q.NE = NE;
  q.NT = NT;
  q.pb = gt;
  q.eb = gt + (G + 1);
  q.rb = gt + 2 * (G + 1);
  q.bb = gt + 3 * (G + 1);
  q.ab = gt + 4 * (G + 1);
  q.tb = gt + 5 * (G + 1);
  const int64_t* d_srcv = gt + 6 * (G + 1);
  const int64_t* d_srca = gt + 7 * (G + 1);
  const int64_t* d_slot = gt + 8 * (G + 1);
  float* P;
  uint64_t* A;
  int32_t *deg, *updeg;
  RCGROW(gpts, float, NV * pdim, P);
  RCGROW(gA, uint64_t, ab[G], A);
  RCGROW(gdeg, int32_t, NV, deg);
  RCGROW(gup, int32_t, NV, updeg);
  hipLaunchKernelGGL(k_rc_gather, dim3(grid_of(NV)), dim3(kBlock), 0, s, q, pdim, d_srcv, d_srca, d_pts,
                     (const uint64_t*)S->A.p, (const int32_t*)S->deg.p, (const int32_t*)S->updeg.p, P, A, deg, updeg);
  const int64_t MAXE = std::max<int64_t>(NE, 1);

  // ---- edges: lexicographic fill, then one stable radix sort by (cloud, length) ----
  int64_t *eoff, *lens, *colptr, *aoff, *cpoff, *cnt, *TB;
  int32_t *erow, *enbr, *vals, *vals2, *ecl, *R, *rstart, *runincl, *runpos, *rlo, *rhi, *inc_r, *epos, *cof, *na,
      *nb, *dims;
  uint64_t *keys, *keys2, *CP, *Bq;
  uint32_t* suv;
  uint16_t *inc_n, *Bp;
  RCGROW(eoff, int64_t, NV + 1, eoff);
  RCGROW(erow, int32_t, MAXE, erow);
  RCGROW(enbr, int32_t, MAXE, enbr);
  RCGROW(keys, uint64_t, MAXE, keys);
  RCGROW(keys2, uint64_t, MAXE, keys2);
  RCGROW(vals, int32_t, MAXE, vals);
  RCGROW(vals2, int32_t, MAXE, vals2);
  if (int rc = scan64(B, s, updeg, NV, eoff, err)) return rc;
  const int gv = grid_of(NV, kWaves);
  hipLaunchKernelGGL(k_rc_efill, dim3(gv), dim3(kBlock), 0, s, q, P, pdim, A, (const int64_t*)eoff, erow, enbr, keys,
                     vals);
  int gbits = 1;
  while ((1ll << gbits) < G) gbits++;
  if (NE > 0) {
    size_t tb2 = 0;
    RCCHK(ssq_sort::sort_pairs<uint64_t>(nullptr, tb2, keys, keys2, vals, vals2, NE, 31 + gbits, s));
    uint8_t* tmp;
    RCGROW(temp, uint8_t, tb2, tmp);
    RCCHK(ssq_sort::sort_pairs<uint64_t>(tmp, tb2, keys, keys2, vals, vals2, NE, 31 + gbits, s));
  }

  // ---- engine buffers (rows are written straight into its pool) ----
  RCGROW(colptr, int64_t, NC + 1, colptr);
  RCGROW(dims, int32_t, NC, dims);
  const int64_t arena = std::max<int64_t>({*arena_hint, NNZ - 2 * NE + 4096, (int64_t)1 << 22});
  int32_t* pool;
  uint32_t* leftone;
  uint8_t* negflag;
  if (int rc = eng->pre_begin(N, NC, NNZ, arena, colptr, dims, 2, &pool, &leftone, &negflag)) {
    if (err) *err = eng->pre_error();
    return rc;
  }
  int32_t* stg = pool + NNZ;  // the arena holds the coboundary lists until the merge

  // ---- ranked edges, rank matrices, runs of equal length ----
  RCGROW(suv, uint32_t, MAXE, suv);
  RCGROW(ecl, int32_t, MAXE, ecl);
  RCGROW(R, int32_t, std::max<int64_t>(rb[G], 1), R);
  RCGROW(rstart, int32_t, MAXE, rstart);
  RCGROW(runincl, int32_t, MAXE, runincl);
  RCGROW(runpos, int32_t, MAXE, runpos);
  RCGROW(rlo, int32_t, MAXE, rlo);
  RCGROW(rhi, int32_t, MAXE, rhi);
  const int ge = grid_of(NE);
  hipLaunchKernelGGL(k_rc_fill32, dim3(grid_of(NV)), dim3(kBlock), 0, s, NV, leftone, kEmpty);  // vertex rows
  hipLaunchKernelGGL(k_rc_dims, dim3(grid_of(NC)), dim3(kBlock), 0, s, NV, NC, dims);
  if (rb[G] > 0)
    hipLaunchKernelGGL(k_rc_fill32, dim3(grid_of(rb[G])), dim3(kBlock), 0, s, rb[G], (uint32_t*)R, 0xffffffffu);
  if (NE > 0) {
    hipLaunchKernelGGL(k_rc_edges, dim3(ge), dim3(kBlock), 0, s, q, (const int32_t*)vals2, (const int32_t*)erow,
                       (const int32_t*)enbr, (const uint64_t*)keys2, suv, ecl, R, rstart, leftone, negflag);
    uint8_t* tmp;
    RCGROW(temp, uint8_t, ssq_sort::scan_bytes(NE), tmp);
    RCCHK((ssq_sort::scan<int32_t, int32_t, true>(rstart, NE, runincl, tmp, s)));
    hipLaunchKernelGGL(k_rc_runpos, dim3(ge), dim3(kBlock), 0, s, NE, (const int32_t*)rstart,
                       (const int32_t*)runincl, runpos);
    hipLaunchKernelGGL(k_rc_runs, dim3(ge), dim3(kBlock), 0, s, NE, (const int32_t*)runincl, (const int32_t*)runpos,
                       rlo, rhi);
  }

  // ---- per vertex: incident edges by rank, checkpoints, the vertex columns ----
  int64_t cp_words = 0;  // bound on the checkpoint words: per vertex (deg / 64 + 1) * W
  for (int g = 0; g < G; g++) {
    const int64_t n = pb[g + 1] - pb[g], W = (n + 63) / 64;
    cp_words += (2 * (eb[g + 1] - eb[g]) / 64 + n) * W;
  }
  int64_t* cpw;
  RCGROW(lens, int64_t, NC, lens);
  RCGROW(aoff, int64_t, NV + 1, aoff);
  RCGROW(cpoff, int64_t, NV + 1, cpoff);
  RCGROW(cpw, int64_t, NV, cpw);
  hipLaunchKernelGGL(k_rc_vlens, dim3(grid_of(NV)), dim3(kBlock), 0, s, q, deg, lens, cpw);
  if (int rc = scan64(B, s, deg, NV, aoff, err)) return rc;
  if (int rc = scan64(B, s, (const int64_t*)cpw, NV, cpoff, err)) return rc;
  if (int rc = scan64(B, s, (const int64_t*)lens, NV, colptr, err)) return rc;  // the vertex block
  RCGROW(inc_r, int32_t, std::max<int64_t>(2 * NE, 1), inc_r);
  RCGROW(inc_n, uint16_t, std::max<int64_t>(2 * NE, 1), inc_n);
  RCGROW(epos, int32_t, std::max<int64_t>(2 * NE, 1), epos);
  RCGROW(CP, uint64_t, std::max<int64_t>(cp_words, 1), CP);
  if (NV > 0)
    hipLaunchKernelGGL(k_rc_vertex, dim3((unsigned)std::min<int64_t>(NV, 65536)), dim3(kBlock), 0, s, q, A,
                       (const int32_t*)R, (const int64_t*)aoff, (const int64_t*)cpoff, (const int64_t*)colptr, inc_r,
                       inc_n, epos, CP, pool);
  RCCHK(hipGetLastError());

  // ---- Q bitsets, triangle bases, column pointers ----
  RCGROW(B, uint64_t, std::max<int64_t>(bb[G], 1), Bq);
  RCGROW(Bp, uint16_t, std::max<int64_t>(bb[G], 1), Bp);
  RCGROW(cnt, int64_t, MAXE, cnt);
  RCGROW(cof, int32_t, MAXE, cof);
  RCGROW(TB, int64_t, NE + 1, TB);
  if (NE > 0)
    hipLaunchKernelGGL(k_rc_qbits, dim3(grid_of(NE, kWaves)), dim3(kBlock), 0, s, q, (const uint32_t*)suv,
                       (const int32_t*)ecl, (const int32_t*)epos, (const int64_t*)aoff, (const int64_t*)cpoff,
                       (const uint16_t*)inc_n, (const uint64_t*)CP, A, Bq, Bp, cnt, cof, lens, negflag);
  if (int rc = scan64(B, s, (const int64_t*)cnt, NE, TB, err)) return rc;
  if (int rc = scan64(B, s, (const int64_t*)lens, NC, colptr, err)) return rc;

  // ---- runs of several equal-length edges: their triangles' lexicographic ranks ----
  int64_t *tcnt, *toff, *tidx;
  int32_t* tiestart;
  RCGROW(tcnt, int64_t, MAXE, tcnt);
  RCGROW(toff, int64_t, NE + 1, toff);
  RCGROW(tiestart, int32_t, MAXE, tiestart);
  RCGROW(tidx, int64_t, NE + 1, tidx);
  int32_t* remap = nullptr;
  int64_t* hx = B->host;
  if (NE > 0) {
    hipLaunchKernelGGL(k_rc_tiecnt, dim3(ge), dim3(kBlock), 0, s, NE, (const int32_t*)rlo, (const int32_t*)rhi,
                       (const int64_t*)cnt, tcnt, tiestart);
    if (int rc = scan64(B, s, (const int64_t*)tcnt, NE, toff, err)) return rc;
    if (int rc = scan64(B, s, (const int32_t*)tiestart, NE, tidx, err)) return rc;
  }
  RCCHK(hipMemcpyAsync(&hx[1], toff + NE, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  RCCHK(hipMemcpyAsync(&hx[2], tidx + NE, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  RCCHK(hipMemcpyAsync(&hx[3], TB + NE, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  RCCHK(hipStreamSynchronize(s));
  const int64_t ntie = NE > 0 ? hx[1] : 0, nrun = NE > 0 ? hx[2] : 0;
  if (hx[3] != NT) {
    if (err) *err = "triangle count mismatch between the sizing pass and the Q bitsets";
    return SSQ_EDEVICE;
  }
  if (ntie > 0) {
    uint64_t *tk, *tk2;
    int32_t *tv, *tv2;
    int64_t *rto, *rr0;
    RCGROW(tkeys, uint64_t, ntie, tk);
    RCGROW(tkeys2, uint64_t, ntie, tk2);
    RCGROW(tvals, int32_t, ntie, tv);
    RCGROW(tvals2, int32_t, ntie, tv2);
    RCGROW(run_toff, int64_t, nrun + 1, rto);
    RCGROW(run_rank0, int64_t, nrun + 1, rr0);
    RCGROW(remap, int32_t, std::max<int64_t>(NT, 1), remap);
    hipLaunchKernelGGL(k_rc_tieenum, dim3(grid_of(NE, kWaves)), dim3(kBlock), 0, s, q, (const uint32_t*)suv,
                       (const int32_t*)ecl, (const int32_t*)rlo, (const int32_t*)rhi, (const int64_t*)tidx,
                       (const int64_t*)toff, (const int64_t*)TB, (const uint64_t*)Bq, (const uint16_t*)Bp, tk, tv);
    hipLaunchKernelGGL(k_rc_tieruns, dim3(ge), dim3(kBlock), 0, s, NE, (const int32_t*)tiestart,
                       (const int64_t*)tidx, (const int64_t*)toff, (const int64_t*)TB, (const int32_t*)ecl, q.tb, rto,
                       rr0);
    int rbits = 1;
    while ((1ll << rbits) < nrun) rbits++;
    size_t tb2 = 0;
    RCCHK(ssq_sort::sort_pairs<uint64_t>(nullptr, tb2, tk, tk2, tv, tv2, ntie, 36 + rbits, s));
    uint8_t* tmp;
    RCGROW(temp, uint8_t, tb2, tmp);
    RCCHK(ssq_sort::sort_pairs<uint64_t>(tmp, tb2, tk, tk2, tv, tv2, ntie, 36 + rbits, s));
    hipLaunchKernelGGL(k_rc_tieapply, dim3(grid_of(ntie)), dim3(kBlock), 0, s, ntie, (const uint64_t*)tk2,
                       (const int32_t*)tv2, (const int64_t*)rto, (const int64_t*)rr0, remap);
  }

  // ---- coboundary columns ----
  RCGROW(na, int32_t, MAXE, na);
  RCGROW(nb, int32_t, MAXE, nb);
  uint32_t* derr;
  RCGROW(err, uint32_t, 4, derr);
  RCCHK(hipMemsetAsync(na, 0, sizeof(int32_t) * MAXE, s));
  RCCHK(hipMemsetAsync(nb, 0, sizeof(int32_t) * MAXE, s));
  RCCHK(hipMemsetAsync(derr, 0, sizeof(uint32_t), s));
  if (NE > 0) {
    hipLaunchKernelGGL(k_rc_side, dim3((unsigned)std::min<int64_t>(NV, 65536)), dim3(kBlock), 0, s, q,
                       (const int64_t*)aoff, (const int32_t*)inc_r, (const uint16_t*)inc_n, (const uint64_t*)Bq,
                       (const uint16_t*)Bp, (const int64_t*)TB, (const int32_t*)rlo, (const int32_t*)rhi,
                       (const int32_t*)remap, (const int64_t*)colptr, (const int64_t*)cnt, (const int32_t*)cof, stg,
                       na, nb);
    hipLaunchKernelGGL(k_rc_merge, dim3(grid_of(NE, kWaves)), dim3(kBlock), 0, s, q, (const int32_t*)ecl,
                       (const int32_t*)rlo, (const int32_t*)rhi, (const int64_t*)TB, (const int32_t*)remap,
                       (const int64_t*)colptr, (const int64_t*)cnt, (const int32_t*)cof, (const int32_t*)na,
                       (const int32_t*)nb, (const int32_t*)stg, pool, leftone, derr);
  }
  hipLaunchKernelGGL(k_rc_check, dim3(grid_of(NC, kWaves)), dim3(kBlock), 0, s, NC, N, (const int64_t*)colptr,
                     (const int32_t*)pool, derr);
  RCCHK(hipGetLastError());
  RCCHK(hipMemcpyAsync(&B->host[4], derr, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  RCCHK(hipStreamSynchronize(s));
  if (const uint32_t bad = *(uint32_t*)&B->host[4]) {
    if (err)
      *err = (bad & 1u) ? "coboundary lists inconsistent with the common-neighbour counts"
                        : "coboundary column not strictly increasing or out of range";
    return SSQ_EDEVICE;
  }

  auto tb1 = std::chrono::steady_clock::now();
  // ---- reduction ----
  ssq_stats st{};
  bool oom = false;
  if (int rc = eng->pre_reduce(prm, &st, &oom)) {
    if (err) *err = eng->pre_error();
    return rc;
  }
  if (oom) {
    *arena_hint = arena * 2;
    return -1;  // rebuild with a larger arena
  }
  acc->n_pending += st.n_pending;
  acc->n_apparent += st.n_apparent;
  acc->n_pages += st.n_pages;
  acc->n_rounds += st.n_rounds;
  acc->n_additions += st.n_additions;
  acc->add_bytes += st.add_bytes;
  acc->t_add_kernel_ms += st.t_add_kernel_ms;
  acc->n_add_launches += st.n_add_launches;
  acc->n_tiled_additions += st.n_tiled_additions;
  acc->n_dense_additions += st.n_dense_additions;
  acc->dense_bytes += st.dense_bytes;
  acc->t_add_long_kernel_ms += st.t_add_long_kernel_ms;
  acc->n_huge_additions += st.n_huge_additions;
  acc->huge_bytes += st.huge_bytes;
  acc->t_add_huge_kernel_ms += st.t_add_huge_kernel_ms;
  acc->max_chain = std::max(acc->max_chain, st.max_chain);
  acc->max_addition_len = std::max(acc->max_addition_len, st.max_addition_len);
  acc->store_bytes = std::max(acc->store_bytes, st.store_bytes);
  acc->t_reduce_ms += st.t_reduce_ms;

  // ---- pairs ----
  int32_t* flag;
  int64_t *pos, *hv, *he;
  RCGROW(flag, int32_t, NC, flag);
  RCGROW(pos, int64_t, NC + 1, pos);
  RCGROW(hv, int64_t, G, hv);
  RCGROW(he, int64_t, G, he);
  hipLaunchKernelGGL(k_rc_pairflag, dim3(grid_of(NC)), dim3(kBlock), 0, s, q, (const int32_t*)ecl, eng->pre_low(),
                     eng->pre_piv(), flag);
  if (int rc = scan64(B, s, (const int32_t*)flag, NC, pos, err)) return rc;
  hipLaunchKernelGGL(k_rc_paircount, dim3(grid_of(G)), dim3(kBlock), 0, s, q, (const int64_t*)pos, hv, he);
  hipLaunchKernelGGL(k_rc_pairemit, dim3(grid_of(NC)), dim3(kBlock), 0, s, q, (const int32_t*)ecl, eng->pre_low(),
                     (const int32_t*)flag, (const int64_t*)pos, (const int64_t*)hv, d_slot, (const int32_t*)rlo,
                     (const int32_t*)rhi, (const int64_t*)TB, stage);
  RCCHK(hipGetLastError());
  std::vector<int64_t> hh(2 * (size_t)G);
  RCCHK(hipMemcpyAsync(hh.data(), hv, sizeof(int64_t) * G, hipMemcpyDeviceToHost, s));
  RCCHK(hipMemcpyAsync(hh.data() + G, he, sizeof(int64_t) * G, hipMemcpyDeviceToHost, s));
  RCCHK(hipStreamSynchronize(s));
  for (int g = 0; g < G; g++) counts[clouds[g]] = hh[(size_t)g] + hh[(size_t)(G + g)];
  if (B->trace) {
    auto tb2 = std::chrono::steady_clock::now();
    std::fprintf(stderr,
                 "[ssq] group of %d clouds (first %d): %lld vertices, %lld edges, %lld triangles, %lld tie triangles; "
                 "build %.1f ms, reduce %.1f ms (init %.1f, rounds %lld, pages %lld, additions %lld, tiled %lld, "
                 "huge %lld, longest chain %lld, pending %lld), pairs %.1f ms\n",
                 G, clouds[0], (long long)NV, (long long)NE, (long long)NT, (long long)ntie,
                 std::chrono::duration<double, std::milli>(tb1 - tb0).count(), st.t_total_ms, st.t_init_ms,
                 (long long)st.n_rounds, (long long)st.n_pages, (long long)st.n_additions,
                 (long long)st.n_tiled_additions, (long long)st.n_huge_additions, (long long)st.max_chain,
                 (long long)st.n_pending,
                 std::chrono::duration<double, std::milli>(tb2 - tb1).count() - st.t_total_ms);
    std::fflush(stderr);
  }
  return SSQ_OK;
}

int ripsco_reduce(RipsCo* const* Bs, PrebuiltTarget* const* engs, int nstreams, int count, const float* d_pts,
                  const int64_t* pbase, int pdim, const float* radius, const ssq_params& p, int64_t* d_pairs,
                  int64_t pairs_cap, int64_t* pair_off, int64_t mem_budget, ssq_stats* st, std::string* err,
                  const int64_t** d_pairs_owned) {
  if (!Bs || !engs || nstreams < 1 || count < 1 || !pbase || !radius || pdim < 1 || !pair_off ||
      (pbase[count] > 0 && !d_pts) || (!d_pairs && !d_pairs_owned))
    return SSQ_EINVAL;
  RipsCo* B = Bs[0];  // holds the sizing pass, the slots and the output
  hipStream_t s = engs[0]->pre_stream();
  auto T0 = std::chrono::steady_clock::now();
  if (int rc = rc_size(B, s, count, d_pts, pbase, pdim, radius, err)) return rc;
  // every cloud gets a slot of (vertices + edges) pairs, an upper bound of its pair count
  std::vector<int64_t> slot((size_t)count + 1, 0);
  for (int c = 0; c < count; c++) slot[c + 1] = slot[c] + (pbase[c + 1] - pbase[c]) + B->h_edges[(size_t)c];
  const int64_t need = slot[count];
  if (!d_pairs) {  // the builder's own buffer
    RCGROW(pairs, int64_t, 2 * std::max<int64_t>(need, 1), d_pairs);
    pairs_cap = need;
    *d_pairs_owned = d_pairs;
  }
  if (pairs_cap < need) {
    if (err) *err = "pair buffer too small (needs one pair per vertex and edge)";
    return SSQ_ENOMEM;
  }
  int64_t *stage, *dslot, *doff;
  RCGROW(stage, int64_t, 2 * std::max<int64_t>(need, 1), stage);
  RCGROW(slots, int64_t, 2 * ((int64_t)count + 1), dslot);
  doff = dslot + count + 1;
  if (mem_budget <= 0)  // a fixed group size (tests force many groups with a small one)
    if (const char* e = std::getenv("SSQ_RIPS_GROUP_BYTES")) mem_budget = std::atoll(e);
  if (mem_budget <= 0) {
    // fixed at the first call, group size included, so that every call groups alike: a later
    // call on fewer streams still finds the other streams' engines holding their buffers
    if (B->budget <= 0) {
      size_t fr = 0, tot = 0;
      RCCHK(hipMemGetInfo(&fr, &tot));
      B->budget = (int64_t)((double)fr * 0.75) - (int64_t)nstreams * kHugeRegionBytes;
      B->budget_streams = nstreams;
    }
    // the groups in flight together; more streams than at the first call also take the extra
    // streams' device regions out of the budget
    const int extra = std::max(0, nstreams - B->budget_streams);
    mem_budget = std::max<int64_t>(1, (B->budget - (int64_t)extra * kHugeRegionBytes) /
                                          std::max(nstreams, B->budget_streams));
  }
  // Groups: clouds by decreasing triangle count, consecutive ones while the group's estimated
  // bytes fit the budget and its cells stay below kGroupCells.  The largest clouds (whose
  // reductions hold the longest chains of additions) share groups, and the groups run on the
  // streams side by side.
  std::vector<int> order((size_t)count);
  for (int c = 0; c < count; c++) order[(size_t)c] = c;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return B->h_tris[(size_t)x] > B->h_tris[(size_t)y]; });
  std::vector<std::pair<int, int>> groups;  // [begin, end) in order
  for (int i0 = 0; i0 < count;) {
    int i1 = i0;
    int64_t nv = 0, ne = 0, nt = 0, n2 = 0, ew = 0;
    while (i1 < count) {
      const int c = order[(size_t)i1];
      const int64_t n = pbase[c + 1] - pbase[c], E = B->h_edges[(size_t)c], T = B->h_tris[(size_t)c];
      const int64_t W = (n + 63) / 64;
      const double bytes = group_bytes(nv + n, ne + E, nt + T, n2 + n * n, ew + E * W, (nt + T) / 16);
      const bool fits = nv + n + ne + E + nt + T <= kGroupCells && bytes <= (double)mem_budget;
      if (!fits && i1 > i0) break;
      nv += n;
      ne += E;
      nt += T;
      n2 += n * n;
      ew += E * W;
      i1++;
      if (!fits) break;  // one cloud alone: try it anyway
    }
    groups.emplace_back(i0, i1);
    i0 = i1;
  }
  std::vector<int64_t> counts((size_t)count, 0);
  std::vector<ssq_stats> accs((size_t)nstreams);
  std::vector<std::string> errs((size_t)nstreams);
  std::vector<int> rcs((size_t)nstreams, SSQ_OK);
  std::atomic<int> next{0};
  std::atomic<bool> stop{false};
  auto worker = [&](int k) {
    if (hipSetDevice(engs[k]->pre_device()) != hipSuccess) {
      rcs[(size_t)k] = SSQ_EDEVICE;
      stop = true;
      return;
    }
    int64_t arena_hint = 0;
    for (int gi; !stop && (gi = next++) < (int)groups.size();) {
      const int i0 = groups[(size_t)gi].first, i1 = groups[(size_t)gi].second;
      int rc;
      for (int attempt = 0;; attempt++) {
        rc = rc_group(Bs[k], B, engs[k], engs[k]->pre_stream(), order.data() + i0, i1 - i0, d_pts, pdim, p, stage,
                      slot.data(), counts.data(), &arena_hint, &accs[(size_t)k], &errs[(size_t)k]);
        if (rc != -1) break;
        if (attempt > 6) {
          errs[(size_t)k] = "column store exhausted";
          rc = SSQ_ENOMEM;
          break;
        }
      }
      if (rc) {
        rcs[(size_t)k] = rc;
        stop = true;
      }
    }
  };
  std::vector<std::thread> pool;
  for (int k = 1; k < nstreams; k++) pool.emplace_back(worker, k);
  worker(0);
  for (auto& t : pool) t.join();
  for (int k = 0; k < nstreams; k++)
    if (rcs[(size_t)k]) {
      if (err) *err = errs[(size_t)k];
      return rcs[(size_t)k];
    }
  // packed output: cloud c's pairs at [pair_off[c], pair_off[c + 1])
  pair_off[0] = 0;
  for (int c = 0; c < count; c++) pair_off[c + 1] = pair_off[c] + counts[(size_t)c];
  RCCHK(hipMemcpyAsync(dslot, slot.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, s));
  RCCHK(hipMemcpyAsync(doff, pair_off, sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_rc_pack, dim3((unsigned)std::min(count, 4096)), dim3(kBlock), 0, s, count,
                     (const int64_t*)dslot, (const int64_t*)doff, (const int64_t*)stage, d_pairs);
  RCCHK(hipGetLastError());
  RCCHK(hipStreamSynchronize(s));
  if (st) {
    ssq_stats acc{};
    for (const ssq_stats& a : accs) {
      acc.n_pending += a.n_pending;
      acc.n_apparent += a.n_apparent;
      acc.n_pages += a.n_pages;
      acc.n_rounds += a.n_rounds;
      acc.n_additions += a.n_additions;
      acc.add_bytes += a.add_bytes;
      acc.t_add_kernel_ms += a.t_add_kernel_ms;
      acc.n_add_launches += a.n_add_launches;
      acc.n_tiled_additions += a.n_tiled_additions;
      acc.n_dense_additions += a.n_dense_additions;
      acc.dense_bytes += a.dense_bytes;
      acc.t_add_long_kernel_ms += a.t_add_long_kernel_ms;
      acc.n_huge_additions += a.n_huge_additions;
      acc.huge_bytes += a.huge_bytes;
      acc.t_add_huge_kernel_ms += a.t_add_huge_kernel_ms;
      acc.max_chain = std::max(acc.max_chain, a.max_chain);
      acc.max_addition_len = std::max(acc.max_addition_len, a.max_addition_len);
      acc.store_bytes = std::max(acc.store_bytes, a.store_bytes);
      acc.t_reduce_ms += a.t_reduce_ms;
    }
    *st = acc;
    st->n_pages = (int64_t)groups.size();  // groups of this call
    st->t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - T0).count();
    st->t_init_ms = 0.0;
  }
  return SSQ_OK;
}

}  // namespace ssq
