// Vietoris-Rips boundary matrices from point clouds (host builder of libssq).
//
// Semantics follow phat-highclear/addons/rips.cpp:49-87: an edge joins two points whose
// distance is <= r; every clique up to maxdim+1 vertices is a simplex; filtration of a
// simplex is its longest edge.  Cells are ordered by (filtration, dimension, lexicographic
// vertex ids).  Distances are float32: sqrt of the float64 sum of squared coordinate
// differences (summed in coordinate order), rounded once to float32.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "ssq_internal.h"

namespace {

struct Cell {
  float f;
  int32_t dim;
  int32_t v[4];
};

inline bool cell_less(const Cell& a, const Cell& b) {
  if (a.f != b.f) return a.f < b.f;
  if (a.dim != b.dim) return a.dim < b.dim;
  for (int k = 0; k <= a.dim; k++)
    if (a.v[k] != b.v[k]) return a.v[k] < b.v[k];
  return false;
}

}  // namespace

namespace ssq {

int vr_build(const float* pts, int64_t npts, int pdim, float r, int maxdim, ssq_matrix* out) {
  std::memset(out, 0, sizeof(*out));
  if (npts < 0 || pdim < 1 || maxdim < 0 || maxdim > 3 || (npts > 0 && !pts) || npts >= (1 << 30))
    return SSQ_EINVAL;
  const int64_t n = npts;
  const int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  // neighbour lists (j > i) with distances
  std::vector<std::vector<std::pair<int32_t, float>>> up((size_t)n);
  auto dist = [&](int64_t i, int64_t j) {
    double acc = 0.0;
    for (int k = 0; k < pdim; k++) {
      const double d = (double)pts[i * pdim + k] - (double)pts[j * pdim + k];
      acc += d * d;
    }
    return (float)std::sqrt(acc);
  };
  if (maxdim >= 1) {
    std::atomic<int64_t> next{0};
    auto work = [&]() {
      for (;;) {
        const int64_t i = next.fetch_add(1);
        if (i >= n) break;
        for (int64_t j = i + 1; j < n; j++) {
          const float d = dist(i, j);
          if (d <= r) up[(size_t)i].emplace_back((int32_t)j, d);
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
  }
  auto edge_f = [&](int32_t a, int32_t b, float* f) -> bool {  // a < b
    const auto& L = up[(size_t)a];
    auto it = std::lower_bound(L.begin(), L.end(), std::make_pair(b, -1.0f),
                               [](const std::pair<int32_t, float>& x, const std::pair<int32_t, float>& y) {
                                 return x.first < y.first;
                               });
    if (it == L.end() || it->first != b) return false;
    *f = it->second;
    return true;
  };
  // enumerate cells
  std::vector<Cell> cells;
  cells.reserve((size_t)n);
  for (int64_t i = 0; i < n; i++) cells.push_back(Cell{0.0f, 0, {(int32_t)i, 0, 0, 0}});
  std::vector<std::vector<Cell>> part((size_t)threads);
  {
    std::atomic<int64_t> next{0};
    auto work = [&](int t) {
      auto& P = part[(size_t)t];
      for (;;) {
        const int64_t i = next.fetch_add(1);
        if (i >= n) break;
        const auto& Li = up[(size_t)i];
        for (size_t a = 0; a < Li.size(); a++) {
          const int32_t j = Li[a].first;
          const float fij = Li[a].second;
          if (maxdim >= 1) P.push_back(Cell{fij, 1, {(int32_t)i, j, 0, 0}});
          if (maxdim < 2) continue;
          for (size_t b = a + 1; b < Li.size(); b++) {
            const int32_t k = Li[b].first;
            float fjk;
            if (!edge_f(j, k, &fjk)) continue;
            const float f3 = std::max(fij, std::max(Li[b].second, fjk));
            P.push_back(Cell{f3, 2, {(int32_t)i, j, k, 0}});
            if (maxdim < 3) continue;
            for (size_t c = b + 1; c < Li.size(); c++) {
              const int32_t l = Li[c].first;
              float fjl, fkl;
              if (!edge_f(j, l, &fjl) || !edge_f(k, l, &fkl)) continue;
              const float f4 = std::max(f3, std::max(Li[c].second, std::max(fjl, fkl)));
              P.push_back(Cell{f4, 3, {(int32_t)i, j, k, l}});
            }
          }
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work, t);
    work(0);
    for (auto& t : pool) t.join();
  }
  for (auto& P : part) {
    cells.insert(cells.end(), P.begin(), P.end());
    std::vector<Cell>().swap(P);
  }
  if ((int64_t)cells.size() >= 0x7fffffffLL) return SSQ_EINVAL;
  std::sort(cells.begin(), cells.end(), cell_less);
  const int64_t N = (int64_t)cells.size();
  // cell index lookups: edges by (i, j) through the neighbour lists; triangles by hash
  std::vector<std::vector<int32_t>> eidx((size_t)n);
  for (int64_t i = 0; i < n; i++) eidx[(size_t)i].assign(up[(size_t)i].size(), -1);
  std::vector<std::pair<uint64_t, int32_t>> tri;
  for (int64_t c = 0; c < N; c++) {
    const Cell& x = cells[(size_t)c];
    if (x.dim == 1) {
      const auto& L = up[(size_t)x.v[0]];
      auto it = std::lower_bound(L.begin(), L.end(), std::make_pair(x.v[1], -1.0f),
                                 [](const std::pair<int32_t, float>& p, const std::pair<int32_t, float>& q) {
                                   return p.first < q.first;
                                 });
      eidx[(size_t)x.v[0]][(size_t)(it - L.begin())] = (int32_t)c;
    } else if (x.dim == 2 && maxdim >= 3) {
      const uint64_t key = ((uint64_t)x.v[0] << 42) | ((uint64_t)x.v[1] << 21) | (uint64_t)x.v[2];
      tri.emplace_back(key, (int32_t)c);
    }
  }
  std::sort(tri.begin(), tri.end());
  auto edge_id = [&](int32_t a, int32_t b) -> int32_t {
    const auto& L = up[(size_t)a];
    auto it = std::lower_bound(L.begin(), L.end(), std::make_pair(b, -1.0f),
                               [](const std::pair<int32_t, float>& p, const std::pair<int32_t, float>& q) {
                                 return p.first < q.first;
                               });
    return eidx[(size_t)a][(size_t)(it - L.begin())];
  };
  auto tri_id = [&](int32_t a, int32_t b, int32_t c) -> int32_t {
    const uint64_t key = ((uint64_t)a << 42) | ((uint64_t)b << 21) | (uint64_t)c;
    auto it = std::lower_bound(tri.begin(), tri.end(), std::make_pair(key, (int32_t)-1));
    return it->second;
  };
  if (maxdim >= 3 && n >= (1 << 21)) return SSQ_EINVAL;
  std::vector<int64_t> colptr((size_t)N + 1, 0);
  std::vector<int32_t> dims((size_t)N);
  for (int64_t c = 0; c < N; c++) {
    dims[(size_t)c] = cells[(size_t)c].dim;
    colptr[(size_t)c + 1] = colptr[(size_t)c] + (cells[(size_t)c].dim ? cells[(size_t)c].dim + 1 : 0);
  }
  std::vector<int32_t> rows((size_t)colptr[(size_t)N]);
  std::vector<int32_t> vid((size_t)n);  // vertex -> cell index (vertices come first)
  for (int64_t c = 0; c < N; c++)
    if (cells[(size_t)c].dim == 0) vid[(size_t)cells[(size_t)c].v[0]] = (int32_t)c;
  for (int64_t c = 0; c < N; c++) {
    const Cell& x = cells[(size_t)c];
    int32_t* o = rows.data() + colptr[(size_t)c];
    if (x.dim == 1) {
      o[0] = vid[(size_t)x.v[0]];
      o[1] = vid[(size_t)x.v[1]];
    } else if (x.dim == 2) {
      o[0] = edge_id(x.v[0], x.v[1]);
      o[1] = edge_id(x.v[0], x.v[2]);
      o[2] = edge_id(x.v[1], x.v[2]);
    } else if (x.dim == 3) {
      o[0] = tri_id(x.v[0], x.v[1], x.v[2]);
      o[1] = tri_id(x.v[0], x.v[1], x.v[3]);
      o[2] = tri_id(x.v[0], x.v[2], x.v[3]);
      o[3] = tri_id(x.v[1], x.v[2], x.v[3]);
    }
    std::sort(o, o + (x.dim ? x.dim + 1 : 0));
  }
  return matrix_from_vectors(colptr, rows, dims, out);
}

}  // namespace ssq

namespace ssq {

// Edges and triangles of the Rips complex at threshold r (same distances as vr_build), by
// adjacency bitsets: triangles are the common neighbours of each edge, counted three times.
// Used to admit point-cloud requests by their real size.
void rips_count_host(const float* pts, int64_t npts, int pdim, float r, int64_t* edges, int64_t* triangles) {
  const int64_t n = npts, W = (n + 63) / 64;
  std::vector<uint64_t> A((size_t)(n * W), 0);
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = i + 1; j < n; j++) {
      double acc = 0.0;
      for (int k = 0; k < pdim; k++) {
        const double d = (double)pts[i * pdim + k] - (double)pts[j * pdim + k];
        acc += d * d;
      }
      if ((float)std::sqrt(acc) <= r) {
        A[(size_t)(i * W + (j >> 6))] |= 1ull << (j & 63);
        A[(size_t)(j * W + (i >> 6))] |= 1ull << (i & 63);
      }
    }
  int64_t e = 0, t3 = 0;
  for (int64_t i = 0; i < n; i++)
    for (int64_t w = i >> 6; w < W; w++) {
      uint64_t m = A[(size_t)(i * W + w)];
      if (w == (i >> 6)) m &= ~((2ull << (i & 63)) - 1);  // neighbours above i
      while (m) {
        const int64_t j = w * 64 + __builtin_ctzll(m);
        m &= m - 1;
        e++;
        for (int64_t x = 0; x < W; x++) t3 += __builtin_popcountll(A[(size_t)(i * W + x)] & A[(size_t)(j * W + x)]);
      }
    }
  *edges = e;
  *triangles = t3 / 3;
}

}  // namespace ssq

extern "C" int ssq_vr_build(const float* pts, int64_t npts, int pdim, float r, int maxdim, ssq_matrix* out) {
  if (!out) return SSQ_EINVAL;
  return ssq::vr_build(pts, npts, pdim, r, maxdim, out);
}
