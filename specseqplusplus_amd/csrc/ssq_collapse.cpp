// Flag-complex edge collapse (model bit 1) — host part of libssq.
//
// Produces exactly the kept-edge set of GUDHI's sequential collapser as SpecSeq++ calls
// it (specseq.h:209-212,410-424 -> Flag_complex_edge_collapser.h:416-482; edge i has
// filtration i).  The algorithm is split in two:
//
//  1. Main-loop decisions.  When the sequential loop tests edge i, every critical edge has
//     index < i, so "valid" (:148-152) reduces to "index < i": the test is a pure function
//     of the graph of edges < i.  These tests are independent and run on all host threads.
//  2. Cascades (set_edge_critical, :326-359).  For each edge that the main loop finds
//     critical, in index order, the non-critical edges of its 3-cliques are re-tested in
//     decreasing index order with the critical set of that moment.  This part is
//     inherently sequential and runs on one thread.  A re-test is first decided from the
//     edge's last witness: the valid edges only grow, so the common neighbours of an edge
//     only grow and its witness's valid neighbourhood only grows; every new common neighbour
//     comes from an edge turned critical, whose 3-clique pass records it for the re-tested
//     edge.  If the witness is adjacent by a valid edge to every recorded vertex that is now a
//     valid common neighbour, the edge is still dominated; otherwise the full test runs.
//
// The domination predicate (:224-303) is: with I the common valid neighbours of (a,b),
// the edge is dominated iff |I| == 1 or some w in I has I inside its closed valid
// neighbourhood.  It does not depend on vertex labels, so vertices are compacted.
#include <algorithm>
#include <atomic>
#include <memory>
#include <queue>
#include <thread>
#include <vector>

#include "ssq_internal.h"

namespace {

struct Graph {
  std::vector<int64_t> off;
  std::vector<int32_t> nbr, eid;
  std::vector<uint8_t> crit;
  // Per edge e = (a, b): its common neighbours x in the whole graph, ascending, with the ids
  // of (a, x) and (b, x) -- the edge's triangles.  The valid common neighbours of a test are
  // a filter of this list, not a merge of two adjacency lists.  Empty when too large.
  std::vector<int64_t> toff;
  int32_t *tx = nullptr, *tea = nullptr, *teb = nullptr;  // edge e's triangles: [toff[e], toff[e] + tcnt[e])
  std::vector<int32_t> tcnt;
  std::vector<int32_t> eu, ev;
  // with at most kBitsetVertices vertices: each neighbourhood as a bit set (W words per
  // vertex) and per word the number of neighbours below it
  std::vector<uint64_t> nb;
  std::vector<int32_t> pre;
  int64_t W = 0;
};
// Storage of the triangle lists (three arrays of int32), grown as needed and never shrunk.
struct TriBuffers {
  std::unique_ptr<int32_t[]> x, a, b;
  size_t cap = 0;
  void reserve(size_t n) {
    if (n <= cap) return;
    const size_t want = n + n / 4;
    x.reset(new int32_t[want]);  // filled on all threads; a vector would zero them on one
    a.reset(new int32_t[want]);
    b.reset(new int32_t[want]);
    cap = want;
  }
};
constexpr int64_t kTriListMax = int64_t(1) << 28;  // entries (12 bytes each)
constexpr int64_t kBitsetVertices = 16384;          // neighbourhood bit sets up to 32 MB

template <typename Valid>
int64_t common(const Graph& g, int32_t a, int32_t b, Valid valid, int32_t* out, int32_t* ea,
               int32_t* eb);

// Valid common neighbours of edge e (ascending), with the two edge ids when ea/eb are given.
template <typename Valid>
int64_t common_e(const Graph& g, int32_t e, Valid valid, int32_t* out, int32_t* ea, int32_t* eb) {
  if (g.toff.empty()) return common(g, g.eu[e], g.ev[e], valid, out, ea, eb);
  int64_t k = 0;
  for (int64_t t = g.toff[e], te = t + g.tcnt[e]; t < te; t++) {
    const int32_t e1 = g.tea[t], e2 = g.teb[t];
    if (!valid(e1) || !valid(e2)) continue;
    if (ea) {
      ea[k] = e1;
      eb[k] = e2;
    }
    out[k++] = g.tx[t];
  }
  return k;
}

template <typename Valid>
int64_t common(const Graph& g, int32_t a, int32_t b, Valid valid, int32_t* out, int32_t* ea,
               int32_t* eb) {
  int64_t i = g.off[a], ie = g.off[a + 1], j = g.off[b], je = g.off[b + 1], k = 0;
  while (i < ie && j < je) {
    if (!valid(g.eid[i])) { i++; continue; }
    if (!valid(g.eid[j])) { j++; continue; }
    int32_t x = g.nbr[i], y = g.nbr[j];
    if (x < y) i++;
    else if (x > y) j++;
    else {
      if (ea) { ea[k] = g.eid[i]; eb[k] = g.eid[j]; }
      if (out) out[k] = x;
      k++;
      i++;
      j++;
    }
  }
  return k;
}

// id of edge (u, v), or -1 (neighbours are ascending)
inline int32_t edge_id(const Graph& g, int32_t u, int32_t v) {
  if (g.W) {
    const uint64_t word = g.nb[(size_t)(u * g.W + (v >> 6))];
    const uint64_t bit = 1ull << (v & 63);
    if (!(word & bit)) return -1;
    return g.eid[g.off[u] + g.pre[(size_t)(u * g.W + (v >> 6))] + __builtin_popcountll(word & (bit - 1))];
  }
  const int32_t* b = g.nbr.data() + g.off[u];
  const int32_t* e = g.nbr.data() + g.off[u + 1];
  const int32_t* p = std::lower_bound(b, e, v);
  return (p != e && *p == v) ? g.eid[p - g.nbr.data()] : -1;
}

// does w's closed valid neighbourhood contain all of I[0:m)?
template <typename Valid>
inline bool covers(const Graph& g, int32_t w, const int32_t* I, int64_t m, Valid valid) {
  if (g.W) {
    for (int64_t t = 0; t < m; t++) {
      if (I[t] == w) continue;
      const int32_t e = edge_id(g, w, I[t]);
      if (e < 0 || !valid(e)) return false;
    }
    return true;
  }
  int64_t p = g.off[w], pe = g.off[w + 1];
  for (int64_t t = 0; t < m; t++) {
    const int32_t x = I[t];
    if (x == w) continue;
    while (p < pe && g.nbr[p] < x) p++;
    if (!(p < pe && g.nbr[p] == x && valid(g.eid[p]))) return false;
  }
  return true;
}

// The predicate of edge_is_dominated: |I| == 1 or some w in I covers I.  *witness is tried
// first and receives the dominating vertex found (the answer does not depend on the order
// in which candidates are tried).
template <typename Valid>
bool dominated(const Graph& g, int32_t e, Valid valid, int32_t* I, int32_t* witness) {
  const int64_t m = common_e(g, e, valid, I, nullptr, nullptr);
  if (m <= 1) {
    if (m == 1) *witness = I[0];
    return m == 1;
  }
  const int32_t hint = *witness;
  if (hint >= 0 && std::binary_search(I, I + m, hint) && covers(g, hint, I, m, valid)) return true;
  for (int64_t c = 0; c < m; c++) {
    const int32_t w = I[c];
    if (w == hint) continue;
    if (covers(g, w, I, m, valid)) {
      *witness = w;
      return true;
    }
  }
  return false;
}

}  // namespace

namespace ssq {

void collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept, int threads) {
  if (ne <= 0) return;
  // compact vertex labels to 0..nv-1 in increasing label order
  int32_t lmin = u[0], lmax = u[0];
  for (int64_t i = 0; i < ne; i++) {
    lmin = std::min(lmin, std::min(u[i], v[i]));
    lmax = std::max(lmax, std::max(u[i], v[i]));
  }
  std::vector<int32_t> eu((size_t)ne), ev((size_t)ne);
  int64_t nv = 0;
  const int64_t range = (int64_t)lmax - lmin + 1;
  if (range <= 8 * ne + 1024) {  // labels in a small range: a direct table
    std::vector<int32_t> id((size_t)range, -1);
    for (int64_t i = 0; i < ne; i++) id[(size_t)(u[i] - lmin)] = id[(size_t)(v[i] - lmin)] = 0;
    for (int64_t x = 0; x < range; x++)
      if (id[(size_t)x] == 0) id[(size_t)x] = (int32_t)nv++;
    for (int64_t i = 0; i < ne; i++) {
      eu[i] = id[(size_t)(u[i] - lmin)];
      ev[i] = id[(size_t)(v[i] - lmin)];
    }
  } else {
    std::vector<int32_t> lab((size_t)(2 * ne));
    for (int64_t i = 0; i < ne; i++) { lab[2 * i] = u[i]; lab[2 * i + 1] = v[i]; }
    std::sort(lab.begin(), lab.end());
    lab.erase(std::unique(lab.begin(), lab.end()), lab.end());
    nv = (int64_t)lab.size();
    for (int64_t i = 0; i < ne; i++) {
      eu[i] = (int32_t)(std::lower_bound(lab.begin(), lab.end(), u[i]) - lab.begin());
      ev[i] = (int32_t)(std::lower_bound(lab.begin(), lab.end(), v[i]) - lab.begin());
    }
  }
  Graph g;
  g.off.assign((size_t)nv + 1, 0);
  for (int64_t i = 0; i < ne; i++) { g.off[eu[i] + 1]++; g.off[ev[i] + 1]++; }
  for (int64_t i = 0; i < nv; i++) g.off[i + 1] += g.off[i];
  g.nbr.resize((size_t)(2 * ne));
  g.eid.resize((size_t)(2 * ne));
  {
    std::vector<int64_t> fill(g.off.begin(), g.off.end() - 1);
    for (int64_t i = 0; i < ne; i++) {
      int64_t p = fill[eu[i]]++;
      g.nbr[p] = ev[i]; g.eid[p] = (int32_t)i;
      p = fill[ev[i]]++;
      g.nbr[p] = eu[i]; g.eid[p] = (int32_t)i;
    }
  }
  if (threads < 1) threads = 1;
  // body(i) for i in [0, count), chunks of `chunk` taken by all threads
  auto parallel_range = [&](int64_t count, int64_t chunk, auto&& body) {
    std::atomic<int64_t> next{0};
    auto work = [&]() {
      for (;;) {
        const int64_t s0 = next.fetch_add(chunk);
        if (s0 >= count) break;
        for (int64_t i = s0; i < std::min(count, s0 + chunk); i++) body(i);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
  };
  auto parallel_edges = [&](auto&& body) { parallel_range(ne, 1024, body); };
  // each vertex's neighbours ascending (they were filled in edge order)
  parallel_range(nv, 16, [&](int64_t a) {
    thread_local std::vector<std::pair<int32_t, int32_t>> tmp;
    tmp.clear();
    for (int64_t p = g.off[a]; p < g.off[a + 1]; p++) tmp.emplace_back(g.nbr[p], g.eid[p]);
    std::sort(tmp.begin(), tmp.end());
    for (int64_t p = g.off[a], k = 0; p < g.off[a + 1]; p++, k++) { g.nbr[p] = tmp[k].first; g.eid[p] = tmp[k].second; }
  });
  int64_t maxdeg = 1;
  for (int64_t a = 0; a < nv; a++) maxdeg = std::max(maxdeg, g.off[a + 1] - g.off[a]);
  g.eu = eu;
  g.ev = ev;
  {  // triangle lists of every edge
    std::vector<int64_t> cap((size_t)ne + 1, 0);
    g.tcnt.assign((size_t)ne, 0);
    const int64_t W = (nv + 63) / 64;
    std::vector<uint64_t>& nb = g.nb;
    std::vector<int32_t>& pre = g.pre;
    if (nv <= kBitsetVertices) {
      g.W = W;
      // neighbourhoods as bit sets: an edge's common neighbours are the set bits of one AND
      // per word, in increasing order, and a neighbour's position in a sorted adjacency list
      // is the popcount of the set bits below it (a prefix count per word plus one word)
      nb.assign((size_t)(nv * W), 0);
      pre.assign((size_t)(nv * W), 0);
      parallel_range(nv, 16, [&](int64_t a) {
        uint64_t* row = nb.data() + a * W;
        for (int64_t p = g.off[a]; p < g.off[a + 1]; p++) row[g.nbr[p] >> 6] |= 1ull << (g.nbr[p] & 63);
        int32_t run = 0;
        for (int64_t w = 0; w < W; w++) {
          pre[(size_t)(a * W + w)] = run;
          run += __builtin_popcountll(row[w]);
        }
      });
      parallel_edges([&](int64_t e) {  // exact counts first
        const uint64_t *ra = nb.data() + (int64_t)eu[e] * W, *rb = nb.data() + (int64_t)ev[e] * W;
        int32_t k = 0;
        for (int64_t w = 0; w < W; w++) k += __builtin_popcountll(ra[w] & rb[w]);
        g.tcnt[e] = k;
      });
      for (int64_t e = 0; e < ne; e++) cap[e + 1] = cap[e] + g.tcnt[e];
    } else {  // merges: edge e gets min(deg a, deg b) slots
      for (int64_t e = 0; e < ne; e++)
        cap[e + 1] = cap[e] + std::min(g.off[eu[e] + 1] - g.off[eu[e]], g.off[ev[e] + 1] - g.off[ev[e]]);
    }
    if (cap[ne] <= kTriListMax) {
      // kept by the calling thread between calls: fresh pages cost more than filling them
      thread_local TriBuffers keep;
      keep.reserve((size_t)cap[ne]);
      g.tx = keep.x.get();
      g.tea = keep.a.get();
      g.teb = keep.b.get();
      if (nv <= kBitsetVertices) {
        parallel_edges([&](int64_t e) {
          const int32_t a = eu[e], b = ev[e];
          const uint64_t *ra = nb.data() + (int64_t)a * W, *rb = nb.data() + (int64_t)b * W;
          const int32_t *pa = pre.data() + (int64_t)a * W, *pb = pre.data() + (int64_t)b * W;
          int32_t* tx = g.tx + cap[e];
          int32_t* ta = g.tea + cap[e];
          int32_t* tb = g.teb + cap[e];
          int32_t k = 0;
          for (int64_t w = 0; w < W; w++) {
            uint64_t m = ra[w] & rb[w];
            while (m) {
              const int bit = __builtin_ctzll(m);
              const uint64_t below = (1ull << bit) - 1;
              tx[k] = (int32_t)(w * 64 + bit);
              ta[k] = g.eid[g.off[a] + pa[w] + __builtin_popcountll(ra[w] & below)];
              tb[k] = g.eid[g.off[b] + pb[w] + __builtin_popcountll(rb[w] & below)];
              k++;
              m &= m - 1;
            }
          }
        });
      } else {
        parallel_edges([&](int64_t e) {
          const int64_t o = cap[e];
          g.tcnt[e] = (int32_t)common(g, eu[e], ev[e], [](int32_t) { return true; }, g.tx + o, g.tea + o, g.teb + o);
        });
      }
      g.toff.swap(cap);
    }
  }

  // 1. main-loop decisions, in parallel
  std::vector<uint8_t> main_crit((size_t)ne, 0);
  std::vector<int32_t> witness((size_t)ne, -1);
  main_crit[0] = 1;
  {
    std::atomic<int64_t> next{1};
    auto work = [&]() {
      std::vector<int32_t> I((size_t)maxdeg);
      for (;;) {
        int64_t s = next.fetch_add(256);
        if (s >= ne) break;
        int64_t e = std::min(ne, s + 256);
        for (int64_t i = s; i < e; i++) {
          int32_t idx = (int32_t)i;
          main_crit[i] = !dominated(g, idx, [idx](int32_t x) { return x < idx; }, I.data(), &witness[i]);
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
  }

  // 2. sequential cascades
  g.crit.assign((size_t)ne, 0);
  std::vector<int32_t> I((size_t)maxdeg), ea((size_t)maxdeg), eb((size_t)maxdeg);
  std::vector<int32_t> stamp((size_t)ne, -1);
  std::priority_queue<int32_t> heap;
  // vertices that may have become common neighbours of a pushed edge since its last test: a
  // list per edge in a pool that lives for one cascade (every pushed edge is popped in it)
  std::vector<int32_t> head((size_t)ne, -1), head_sid((size_t)ne, -1), pool_next, pool_x;
  auto add_cand = [&](int32_t e, int32_t x, int32_t sid) {
    if (head_sid[e] != sid) {
      head_sid[e] = sid;
      head[e] = -1;
    }
    pool_next.push_back(head[e]);
    pool_x.push_back(x);
    head[e] = (int32_t)pool_x.size() - 1;
  };
  auto three_clique = [&](int32_t thresh, int32_t e, int32_t sid) {
    const uint8_t* crit = g.crit.data();
    int64_t m = common_e(g, e, [thresh, crit](int32_t x) { return x < thresh || crit[x]; }, I.data(), ea.data(),
                         eb.data());
    const int32_t p = g.eu[e], q = g.ev[e];
    for (int64_t k = 0; k < m; k++) {
      const int32_t e1 = ea[k], e2 = eb[k];  // e1 = (p, x) gains q, e2 = (q, x) gains p
      if (!crit[e1] && e1 < thresh) {
        add_cand(e1, q, sid);
        if (stamp[e1] != sid) { stamp[e1] = sid; heap.push(e1); }
      }
      if (!crit[e2] && e2 < thresh) {
        add_cand(e2, p, sid);
        if (stamp[e2] != sid) { stamp[e2] = sid; heap.push(e2); }
      }
    }
  };
  // still dominated by its last witness: every recorded vertex is the witness, or joined to it
  // by a valid edge, or not a valid common neighbour
  auto witness_holds = [&](int32_t cur, int32_t sid) {
    const int32_t w = witness[cur];
    if (w < 0) return false;
    const uint8_t* crit = g.crit.data();
    auto valid = [cur, crit](int32_t x) { return x >= 0 && (x < cur || crit[x]); };
    const int32_t a = g.eu[cur], b = g.ev[cur];
    for (int32_t t = head_sid[cur] == sid ? head[cur] : -1; t >= 0; t = pool_next[t]) {
      const int32_t x = pool_x[t];
      if (x == w || valid(edge_id(g, w, x))) continue;
      if (valid(edge_id(g, a, x)) && valid(edge_id(g, b, x))) return false;
    }
    return true;
  };
  std::fill(kept, kept + ne, 0);
  g.crit[0] = 1;
  kept[0] = 1;
  for (int32_t idx = 1; idx < (int32_t)ne; idx++) {
    if (!main_crit[idx]) continue;
    pool_next.clear();
    pool_x.clear();
    g.crit[idx] = 1;
    kept[idx] = 1;
    three_clique(idx, idx, idx);
    while (!heap.empty()) {
      int32_t cur = heap.top();
      heap.pop();
      if (g.crit[cur]) continue;
      if (witness_holds(cur, idx)) continue;
      const uint8_t* crit = g.crit.data();
      if (dominated(g, cur, [cur, crit](int32_t x) { return x < cur || crit[x]; }, I.data(), &witness[cur]))
        continue;
      g.crit[cur] = 1;
      kept[cur] = 1;
      three_clique(cur, cur, idx);
    }
  }
}

}  // namespace ssq

extern "C" int ssq_collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept, int threads) {
  if (ne < 0 || (ne > 0 && (!u || !v || !kept))) return SSQ_EINVAL;
  if (threads <= 0) threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  ssq::collapse_edges(ne, u, v, kept, threads);
  return SSQ_OK;
}
