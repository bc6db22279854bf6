// Vietoris-Rips boundary matrices built on the GPU (dimension <= 2).
//
// Same complex and cell order as the host builder (ssq_gen.cpp; semantics of
// phat-highclear/addons/rips.cpp:49-87): an edge joins points at float32 distance <= r
// (sqrt of the float64 sum of squared coordinate differences, summed in coordinate order,
// rounded once), triangles are the 3-cliques, the filtration value of a cell is its
// longest edge, and cells are ordered by (filtration, dimension, lexicographic vertices).
//
// Steps: per-row neighbour counts and an ordered fill of the "up" lists (neighbours with a
// larger id), triangles by merging up(i) past j with up(j) for every edge (i,j), then one
// stable radix sort of (float bits << 2 | dim) over cells generated in lexicographic order
// (vertices, edges by (i,j), triangles by (i,j,k)), which yields exactly the
// (filtration, dimension, lex) order, and the CSC boundary assembled from the inverse
// permutation.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ssq_internal.h"
#include "ssq_kernels.h"

namespace ssq {

namespace {

constexpr int kVrBlock = 256;

__device__ __forceinline__ float vr_dist(const float* P, int64_t i, int64_t j, int pdim) {
  return rips_dist(P, i, j, pdim);
}

__global__ void k_vr_set(int64_t* p, int64_t v) { *p = v; }

// Clouds of a batch are concatenated: cloud c owns points [pbase[c], pbase[c + 1]) and has
// threshold rad[c].  Returns the cloud of point i.
__device__ __forceinline__ int64_t vr_cloud(const int64_t* pbase, int64_t count, int64_t i) {
  int64_t lo = 0, hi = count;  // largest c with pbase[c] <= i
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (pbase[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// one workgroup per row i: number of j > i of the same cloud within its threshold
__global__ void k_vr_deg(const float* P, int64_t n, int pdim, const int64_t* pbase, const float* rad, int64_t count,
                         int64_t* deg) {
  __shared__ int32_t s[kVrBlock / kWave];
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int64_t cl = vr_cloud(pbase, count, i), end = pbase[cl + 1];
    const float r = rad[cl];
    int32_t c = 0;
    for (int64_t j = i + 1 + threadIdx.x; j < end; j += kVrBlock) c += vr_dist(P, i, j, pdim) <= r;
    for (int d = kWave / 2; d > 0; d >>= 1) c += __shfl_xor(c, d, kWave);
    if (lane_id() == 0) s[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t t = 0;
      for (int w = 0; w < kVrBlock / kWave; w++) t += s[w];
      deg[i] = t;
    }
    __syncthreads();
  }
}

// ordered fill of the up lists: nbr[eoff[i] ..] = ascending j > i within r
__global__ void k_vr_edges(const float* P, int64_t n, int pdim, const int64_t* pbase, const float* rad, int64_t count,
                           const int64_t* eoff, int32_t* nbr, float* ed, int32_t* erow) {
  __shared__ int32_t s[kVrBlock / kWave];
  const int w = threadIdx.x / kWave, lane = lane_id();
  const uint64_t lower = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int64_t cl = vr_cloud(pbase, count, i), end = pbase[cl + 1];
    const float r = rad[cl];
    int64_t base = eoff[i];
    for (int64_t j0 = i + 1; j0 < end; j0 += kVrBlock) {
      const int64_t j = j0 + threadIdx.x;
      float d = 0.f;
      bool in = false;
      if (j < end) {
        d = vr_dist(P, i, j, pdim);
        in = d <= r;
      }
      const uint64_t m = __ballot(in);
      if (lane == 0) s[w] = __popcll(m);
      __syncthreads();
      int32_t before = 0, total = 0;
      for (int k = 0; k < kVrBlock / kWave; k++) {
        if (k < w) before += s[k];
        total += s[k];
      }
      if (in) {
        const int64_t pos = base + before + __popcll(m & lower);
        nbr[pos] = (int32_t)j;
        ed[pos] = d;
        erow[pos] = (int32_t)i;
      }
      base += total;
      __syncthreads();
    }
  }
}

// triangles per edge (i,j): k in up(i) after j and in up(j)
__global__ void k_vr_tri_count(int64_t ne, const int64_t* eoff, const int32_t* nbr, const int32_t* erow,
                               int64_t* tcnt) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = erow[e], j = nbr[e];
    int64_t a = e + 1, ae = eoff[i + 1], b = eoff[j], be = eoff[j + 1];
    int64_t c = 0;
    while (a < ae && b < be) {
      const int32_t x = nbr[a], y = nbr[b];
      if (x < y) a++;
      else if (y < x) b++;
      else {
        c++;
        a++;
        b++;
      }
    }
    tcnt[e] = c;
  }
}

__global__ void k_vr_tri_fill(int64_t ne, const int64_t* eoff, const int32_t* nbr, const int32_t* erow, const float* ed,
                              const int64_t* toff, int32_t* tedge, float* tf) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = erow[e], j = nbr[e];
    int64_t a = e + 1, ae = eoff[i + 1], b = eoff[j], be = eoff[j + 1];
    int64_t t = toff[e];
    while (a < ae && b < be) {
      const int32_t x = nbr[a], y = nbr[b];
      if (x < y) a++;
      else if (y < x) b++;
      else {
        tedge[3 * t] = (int32_t)e;      // (i,j)
        tedge[3 * t + 1] = (int32_t)a;  // (i,k)
        tedge[3 * t + 2] = (int32_t)b;  // (j,k)
        tf[t] = fmaxf(ed[e], fmaxf(ed[a], ed[b]));
        t++;
        a++;
        b++;
      }
    }
  }
}

// sort keys in generation order (vertices, edges, triangles; cloud-major and lexicographic
// within each): cloud << 34 | float bits << 2 | dim.  Filtration values are >= 0, so their
// bits order like the values.
__global__ void k_vr_keys(int64_t n, int64_t ne, int64_t nt, const float* ed, const float* tf, const int32_t* erow,
                          const int32_t* tedge, const int64_t* pbase, int64_t count, uint64_t* keys, int32_t* vals) {
  const int64_t N = n + ne + nt;
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < N; c += (int64_t)gridDim.x * blockDim.x) {
    uint64_t k = 0;
    int64_t v;  // a vertex of the cell, for its cloud
    if (c < n) v = c;
    else if (c < n + ne) {
      v = erow[c - n];
      k = ((uint64_t)__float_as_uint(ed[c - n]) << 2) | 1u;
    } else {
      v = erow[tedge[3 * (c - n - ne)]];
      k = ((uint64_t)__float_as_uint(tf[c - n - ne]) << 2) | 2u;
    }
    if (count > 1) k |= (uint64_t)vr_cloud(pbase, count, v) << 34;
    keys[c] = k;
    vals[c] = (int32_t)c;
  }
}

// cells of each cloud: points + edges + triangles, from the scans over rows and edges
__global__ void k_vr_cloud_cells(int64_t count, const int64_t* pbase, const int64_t* eoff, const int64_t* toff,
                                 int64_t* cells) {
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < count; c += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e0 = eoff[pbase[c]], e1 = eoff[pbase[c + 1]];
    cells[c] = (pbase[c + 1] - pbase[c]) + (e1 - e0) + (toff ? toff[e1] - toff[e0] : 0);
  }
}

__global__ void k_vr_inverse(int64_t N, const int32_t* order, int32_t* pos, int64_t n, int64_t ne, int32_t* dims,
                             int64_t* cnt) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int32_t id = order[p];
    pos[id] = (int32_t)p;
    const int32_t dm = id < n ? 0 : (id < n + ne ? 1 : 2);
    dims[p] = dm;
    cnt[p] = dm == 0 ? 0 : dm + 1;
  }
}

__global__ void k_vr_rows(int64_t N, const int32_t* order, const int32_t* pos, int64_t n, int64_t ne,
                          const int32_t* erow, const int32_t* nbr, const int32_t* tedge, const int64_t* colptr,
                          int32_t* rows) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
    const int64_t id = order[p];
    int32_t* o = rows + colptr[p];
    if (id < n) continue;
    if (id < n + ne) {  // a cloud's vertices come first in its block, in id order
      o[0] = pos[erow[id - n]];
      o[1] = pos[nbr[id - n]];
    } else {
      const int64_t t = id - n - ne;
      int32_t a = pos[n + tedge[3 * t]], b = pos[n + tedge[3 * t + 1]], c = pos[n + tedge[3 * t + 2]];
      if (a > b) { const int32_t x = a; a = b; b = x; }
      if (b > c) { const int32_t x = b; b = c; c = x; }
      if (a > b) { const int32_t x = a; a = b; b = x; }
      o[0] = a;
      o[1] = b;
      o[2] = c;
    }
  }
}

template <typename T>
bool grow(void*& p, size_t& bytes, size_t count) {
  const size_t need = std::max<size_t>(count, 1) * sizeof(T);
  if (bytes >= need) return true;
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  if (hipMalloc(&p, need + need / 2) != hipSuccess) {  // geometric growth: a free synchronizes the device
    (void)hipGetLastError();
    return false;
  }
  bytes = need + need / 2;
  return true;
}

}  // namespace

struct VrBuilder {
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  Buf pts, deg, eoff, nbr, ed, erow, tcnt, toff, tedge, tf, keys, keys2, vals, vals2, pos, dims, cnt, colptr, rows,
      temp, pbase, rad, cells;
  int64_t host2[2] = {0, 0};
  std::string err;
  ~VrBuilder() {
    for (Buf* b : {&pts, &deg, &eoff, &nbr, &ed, &erow, &tcnt, &toff, &tedge, &tf, &keys, &keys2, &vals, &vals2, &pos,
                   &dims, &cnt, &colptr, &rows, &temp, &pbase, &rad, &cells})
      if (b->p) (void)hipFree(b->p);
  }
};

VrBuilder* vr_builder_create() { return new (std::nothrow) VrBuilder(); }
void vr_builder_destroy(VrBuilder* b) { delete b; }

#define VRCHK(x)                    \
  do {                              \
    if ((x) != hipSuccess) {        \
      (void)hipGetLastError();      \
      return SSQ_EDEVICE;           \
    }                               \
  } while (0)
#define VRGROW(buf, T, n)                                           \
  do {                                                              \
    if (!grow<T>(B->buf.p, B->buf.bytes, (size_t)(n))) return SSQ_ENOMEM; \
  } while (0)

int vr_builder_build_batch(VrBuilder* B, hipStream_t s, int count, const float* const* h_pts, const int64_t* npts,
                           int pdim, const float* radius, int maxdim, const int64_t** d_colptr, const int32_t** d_rows,
                           const int32_t** d_dims, int64_t* n_out, int64_t* nnz_out, int64_t* cell_base) {
  if (!B || count < 1 || pdim < 1 || maxdim < 0 || maxdim > 2 || !h_pts || !npts || !radius || count > (1 << 20))
    return SSQ_EINVAL;
  std::vector<int64_t> pb((size_t)count + 1, 0);
  for (int c = 0; c < count; c++) {
    if (npts[c] < 0 || (npts[c] && !h_pts[c]) || !(radius[c] >= 0.f)) return SSQ_EINVAL;
    pb[c + 1] = pb[c] + npts[c];
  }
  const int64_t n = pb[count];
  if (n >= (1 << 30)) return SSQ_EINVAL;
  VRGROW(pts, float, n * pdim);
  VRGROW(pbase, int64_t, count + 1);
  VRGROW(rad, float, count);
  VRGROW(cells, int64_t, count + 1);
  for (int c = 0; c < count; c++)
    if (npts[c])
      VRCHK(hipMemcpyAsync((float*)B->pts.p + pb[c] * pdim, h_pts[c], sizeof(float) * npts[c] * pdim,
                           hipMemcpyHostToDevice, s));
  VRCHK(hipMemcpyAsync(B->pbase.p, pb.data(), sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, s));
  VRCHK(hipMemcpyAsync(B->rad.p, radius, sizeof(float) * count, hipMemcpyHostToDevice, s));
  const float* P = (const float*)B->pts.p;
  const int64_t* dpb = (const int64_t*)B->pbase.p;
  const float* drad = (const float*)B->rad.p;
  int64_t ne = 0, nt = 0;
  VRGROW(eoff, int64_t, n + 1);
  int64_t* eoff = (int64_t*)B->eoff.p;
  const int64_t* toffp = nullptr;
  if (maxdim >= 1 && n > 1) {
    VRGROW(deg, int64_t, n);
    const int g = (int)std::min<int64_t>(n, 8192);
    hipLaunchKernelGGL(k_vr_deg, dim3(g), dim3(kVrBlock), 0, s, P, n, pdim, dpb, drad, (int64_t)count,
                       (int64_t*)B->deg.p);
    // exclusive scan of degrees (hipcub), edge count from the last entries
    size_t tb = 0;
    VRCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int64_t*)B->deg.p, eoff, (int)n, s));
    VRGROW(temp, uint8_t, tb);
    VRCHK(hipcub::DeviceScan::ExclusiveSum(B->temp.p, tb, (int64_t*)B->deg.p, eoff, (int)n, s));
    int64_t* hx = B->host2;
    VRCHK(hipMemcpyAsync(&hx[0], eoff + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    VRCHK(hipMemcpyAsync(&hx[1], (int64_t*)B->deg.p + n - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    VRCHK(hipStreamSynchronize(s));
    ne = hx[0] + hx[1];
    hipLaunchKernelGGL(k_vr_set, dim3(1), dim3(1), 0, s, eoff + n, ne);
    VRGROW(nbr, int32_t, ne);
    VRGROW(ed, float, ne);
    VRGROW(erow, int32_t, ne);
    hipLaunchKernelGGL(k_vr_edges, dim3(g), dim3(kVrBlock), 0, s, P, n, pdim, dpb, drad, (int64_t)count, eoff,
                       (int32_t*)B->nbr.p, (float*)B->ed.p, (int32_t*)B->erow.p);
    if (maxdim >= 2 && ne > 0) {
      VRGROW(tcnt, int64_t, ne);
      VRGROW(toff, int64_t, ne + 1);
      const int ge = (int)std::min<int64_t>((ne + 255) / 256, 16384);
      hipLaunchKernelGGL(k_vr_tri_count, dim3(ge), dim3(256), 0, s, ne, eoff, (int32_t*)B->nbr.p, (int32_t*)B->erow.p,
                         (int64_t*)B->tcnt.p);
      size_t tb2 = 0;
      VRCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (int64_t*)B->tcnt.p, (int64_t*)B->toff.p, (int)ne, s));
      VRGROW(temp, uint8_t, tb2);
      VRCHK(hipcub::DeviceScan::ExclusiveSum(B->temp.p, tb2, (int64_t*)B->tcnt.p, (int64_t*)B->toff.p, (int)ne, s));
      VRCHK(hipMemcpyAsync(&hx[0], (int64_t*)B->toff.p + ne - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      VRCHK(hipMemcpyAsync(&hx[1], (int64_t*)B->tcnt.p + ne - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
      VRCHK(hipStreamSynchronize(s));
      nt = hx[0] + hx[1];
      if (n + ne + nt >= 0x7fffffffLL) return SSQ_EINVAL;
      hipLaunchKernelGGL(k_vr_set, dim3(1), dim3(1), 0, s, (int64_t*)B->toff.p + ne, nt);
      VRGROW(tedge, int32_t, 3 * nt);
      VRGROW(tf, float, nt);
      hipLaunchKernelGGL(k_vr_tri_fill, dim3(ge), dim3(256), 0, s, ne, eoff, (int32_t*)B->nbr.p, (int32_t*)B->erow.p,
                         (float*)B->ed.p, (int64_t*)B->toff.p, (int32_t*)B->tedge.p, (float*)B->tf.p);
      toffp = (const int64_t*)B->toff.p;
    }
  } else {
    VRCHK(hipMemsetAsync(eoff, 0, sizeof(int64_t) * (n + 1), s));
  }
  // cells per cloud, for the block bases
  hipLaunchKernelGGL(k_vr_cloud_cells, dim3((unsigned)std::min<int64_t>((count + 255) / 256, 1024)), dim3(256), 0, s,
                     (int64_t)count, dpb, eoff, toffp, (int64_t*)B->cells.p);
  const int64_t N = n + ne + nt;
  VRGROW(keys, uint64_t, N);
  VRGROW(keys2, uint64_t, N);
  VRGROW(vals, int32_t, N);
  VRGROW(vals2, int32_t, N);
  VRGROW(pos, int32_t, N);
  VRGROW(dims, int32_t, N);
  VRGROW(cnt, int64_t, N);
  VRGROW(colptr, int64_t, N + 1);
  const int64_t nnz = 2 * ne + 3 * nt;
  VRGROW(rows, int32_t, nnz);
  const int gN = (int)std::max<int64_t>(1, std::min<int64_t>((N + 255) / 256, 16384));
  if (N > 0) {
    hipLaunchKernelGGL(k_vr_keys, dim3(gN), dim3(256), 0, s, n, ne, nt, (float*)B->ed.p, (float*)B->tf.p,
                       (const int32_t*)B->erow.p, (const int32_t*)B->tedge.p, dpb, (int64_t)count,
                       (uint64_t*)B->keys.p, (int32_t*)B->vals.p);
    int cbits = 0;
    while ((1ll << cbits) < count) cbits++;
    const int end_bit = 34 + (count > 1 ? cbits : 0);
    size_t tb = 0;
    VRCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint64_t*)B->keys.p, (uint64_t*)B->keys2.p,
                                             (int32_t*)B->vals.p, (int32_t*)B->vals2.p, (int)N, 0, end_bit, s));
    VRGROW(temp, uint8_t, tb);
    VRCHK(hipcub::DeviceRadixSort::SortPairs(B->temp.p, tb, (uint64_t*)B->keys.p, (uint64_t*)B->keys2.p,
                                             (int32_t*)B->vals.p, (int32_t*)B->vals2.p, (int)N, 0, end_bit, s));
    hipLaunchKernelGGL(k_vr_inverse, dim3(gN), dim3(256), 0, s, N, (int32_t*)B->vals2.p, (int32_t*)B->pos.p, n, ne,
                       (int32_t*)B->dims.p, (int64_t*)B->cnt.p);
    size_t tb3 = 0;
    VRCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, (int64_t*)B->cnt.p, (int64_t*)B->colptr.p, (int)N, s));
    VRGROW(temp, uint8_t, tb3);
    VRCHK(hipcub::DeviceScan::ExclusiveSum(B->temp.p, tb3, (int64_t*)B->cnt.p, (int64_t*)B->colptr.p, (int)N, s));
    hipLaunchKernelGGL(k_vr_set, dim3(1), dim3(1), 0, s, (int64_t*)B->colptr.p + N, nnz);
    hipLaunchKernelGGL(k_vr_rows, dim3(gN), dim3(256), 0, s, N, (int32_t*)B->vals2.p, (int32_t*)B->pos.p, n, ne,
                       (int32_t*)B->erow.p, (int32_t*)B->nbr.p, (int32_t*)B->tedge.p, (int64_t*)B->colptr.p,
                       (int32_t*)B->rows.p);
  } else {
    hipLaunchKernelGGL(k_vr_set, dim3(1), dim3(1), 0, s, (int64_t*)B->colptr.p, (int64_t)0);
  }
  VRCHK(hipGetLastError());
  if (cell_base) {
    std::vector<int64_t> cells((size_t)count);
    VRCHK(hipMemcpyAsync(cells.data(), B->cells.p, sizeof(int64_t) * count, hipMemcpyDeviceToHost, s));
    VRCHK(hipStreamSynchronize(s));
    cell_base[0] = 0;
    for (int c = 0; c < count; c++) cell_base[c + 1] = cell_base[c] + cells[c];
  } else {
    VRCHK(hipStreamSynchronize(s));
  }
  *d_colptr = (const int64_t*)B->colptr.p;
  *d_rows = (const int32_t*)B->rows.p;
  *d_dims = (const int32_t*)B->dims.p;
  *n_out = N;
  *nnz_out = nnz;
  return SSQ_OK;
}

int vr_builder_build(VrBuilder* B, hipStream_t s, const float* h_pts, int64_t npts, int pdim, float r, int maxdim,
                     const int64_t** d_colptr, const int32_t** d_rows, const int32_t** d_dims, int64_t* n_out,
                     int64_t* nnz_out) {
  if (npts < 0 || (npts && !h_pts)) return SSQ_EINVAL;
  return vr_builder_build_batch(B, s, 1, &h_pts, &npts, pdim, &r, maxdim, d_colptr, d_rows, d_dims, n_out, nnz_out,
                                nullptr);
}

}  // namespace ssq
