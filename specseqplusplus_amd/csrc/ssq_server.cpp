// Continuous-batching request server on one GPU (new; no reference counterpart).
//
// Requests (boundary matrices, point clouds that are turned into Vietoris-Rips matrices, or
// matrices registered once and kept resident in device memory) wait in one queue per kind.  Each worker owns an engine handle with its own HIP stream.  A worker
// that is free takes the oldest requests that fit both max_batch and the shared budget of
// estimated device bytes, reduces them as one block-diagonal batch, publishes each result
// and returns the budget.  With several workers, a batch is formed and launched while other
// batches are still on the device, so requests that arrive during a batch do not wait for it
// to finish.  Each result is the pair set of its own matrix: block-diagonal batching never
// mixes rows of different requests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ssq_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id;
  Clock::time_point t_submit;
  std::vector<int64_t> colptr;
  std::vector<int32_t> rows, dims;
  std::vector<float> pts;
  int64_t npts = 0;
  int pdim = 0, maxdim = 0;
  float r = 0;
  bool is_points = false;
  int64_t mid = 0;  // registered matrix (0: none)
  int64_t est_bytes = 0;
  int64_t est_cells = 0;
};

struct Result {
  std::vector<int64_t> pairs, ess;
  double latency_ms = 0;
  int status = SSQ_OK;
};

int64_t estimate_bytes(int64_t n, int64_t nnz) {
  // input copy + column store arena (4x nnz) + per-cell state (~48 B) + pairs
  return 24 * nnz + 64 * n + (1 << 20);
}

// A point cloud's cells, counted exactly on the host at submission (adjacency bitsets: the
// edges within r, the triangles as common neighbours), and its device bytes as a matrix of
// that size.  Dimension 3 is bounded by the tetrahedra of the triangle count's cube root.
void estimate_points(const float* pts, int64_t npts, int pdim, float r, int maxdim, int64_t* cells, int64_t* bytes) {
  int64_t e = 0, t = 0;
  if (maxdim >= 1 && npts > 1) ssq::rips_count_host(pts, npts, pdim, r, &e, &t);
  if (maxdim < 2) t = 0;
  const int64_t tets = maxdim >= 3 ? t * (int64_t)std::cbrt((double)std::max<int64_t>(t, 1)) : 0;
  *cells = npts + e + t + tets;
  *bytes = estimate_bytes(*cells, 2 * e + 3 * t + 4 * tets);
}

// A matrix registered for resident requests: its device copy and the requests that use it.
struct Resident {
  ssq_device_matrix dm{};
  int64_t refs = 0;  // queued or running requests
  bool closing = false;
  int64_t est_bytes = 0;
};

// Host checks of a submitted matrix (rows are range-checked on the device): offsets start at
// 0, never decrease and end at nnz.
bool matrix_ok(const ssq_matrix* m) {
  if (m->n < 0 || m->nnz < 0 || (m->n && (!m->col_ptr || !m->dims)) || (m->nnz && !m->rows)) return false;
  if (m->n == 0) return m->nnz == 0 || !m->col_ptr || m->col_ptr[0] == 0;
  if (m->col_ptr[0] != 0 || m->col_ptr[m->n] != m->nnz) return false;
  for (int64_t c = 0; c < m->n; c++)
    if (m->col_ptr[c + 1] < m->col_ptr[c]) return false;
  return true;
}

// The device-side checks, on the host, for a matrix that is registered once: every column's
// rows strictly increasing and in [0, n).
bool rows_ok(const ssq_matrix* m) {
  for (int64_t c = 0; c < m->n; c++) {
    int64_t prev = -1;
    for (int64_t k = m->col_ptr[c]; k < m->col_ptr[c + 1]; k++) {
      const int64_t r = m->rows[k];
      if (r <= prev || r >= m->n) return false;
      prev = r;
    }
  }
  return true;
}

// Triangulated k x k grid (vertices, then edges, then triangles: a valid filtration) with
// about `cells` cells.  Reducing it sizes a worker's device buffers before requests arrive,
// so that no buffer growth (whose free synchronizes the device) lands in a served batch.
void grid_complex(int64_t cells, std::vector<int64_t>& cp, std::vector<int32_t>& rows, std::vector<int32_t>& dims) {
  int64_t k = 2;
  while (6 * (k + 1) * (k + 1) <= cells) k++;
  const int64_t V = k * k;
  auto vid = [k](int64_t x, int64_t y) { return (int32_t)(y * k + x); };
  std::vector<std::pair<int32_t, int32_t>> E;
  for (int64_t y = 0; y < k; y++)
    for (int64_t x = 0; x < k; x++) {
      if (x + 1 < k) E.push_back({vid(x, y), vid(x + 1, y)});
      if (y + 1 < k) E.push_back({vid(x, y), vid(x, y + 1)});
      if (x + 1 < k && y + 1 < k) E.push_back({vid(x, y), vid(x + 1, y + 1)});
    }
  std::sort(E.begin(), E.end());
  auto eid = [&](int32_t a, int32_t b) {
    return (int32_t)(V + (std::lower_bound(E.begin(), E.end(), std::make_pair(a, b)) - E.begin()));
  };
  rows.clear();
  dims.assign((size_t)V, 0);
  for (auto& e : E) {
    rows.push_back(e.first);
    rows.push_back(e.second);
    dims.push_back(1);
  }
  for (int64_t y = 0; y + 1 < k; y++)
    for (int64_t x = 0; x + 1 < k; x++) {
      const int32_t a = vid(x, y), b = vid(x + 1, y), c = vid(x, y + 1), d = vid(x + 1, y + 1);
      int32_t t1[3] = {eid(a, b), eid(a, d), eid(b, d)}, t2[3] = {eid(a, c), eid(a, d), eid(c, d)};
      for (int32_t* t : {t1, t2}) {
        std::sort(t, t + 3);
        rows.insert(rows.end(), t, t + 3);
        dims.push_back(2);
      }
    }
  cp.assign(dims.size() + 1, 0);
  for (size_t c = 0; c < dims.size(); c++) cp[c + 1] = cp[c] + (dims[c] == 0 ? 0 : dims[c] + 1);
}

void warm_up(ssq_handle* h, const ssq_server_config& cfg) {
  const int64_t cells = std::min<int64_t>(cfg.max_batch_cells > 0 ? cfg.max_batch_cells : 1000000, 4000000);
  std::vector<int64_t> cp;
  std::vector<int32_t> rows, dims;
  grid_complex(cells, cp, rows, dims);
  ssq_matrix m{(int64_t)dims.size(), (int64_t)rows.size(), cp.data(), rows.data(), dims.data()};
  const ssq_matrix* one[1] = {&m};
  for (const ssq_params* p : {&cfg.params, &cfg.points_params}) {
    ssq_pairs out;
    if (ssq_batch_reduce(h, one, 1, p, &out, nullptr) == SSQ_OK) ssq_pairs_free(&out);
  }
  // point clouds: random points in the unit cube, about 20 neighbours each
  constexpr int kClouds = 16, kPts = 2048;
  std::vector<float> pts((size_t)kClouds * kPts * 3);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (float& v : pts) {  // splitmix64
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    v = (float)((z ^ (z >> 31)) >> 40) / (float)(1ull << 24);
  }
  std::vector<const float*> pp(kClouds);
  std::vector<int64_t> pn(kClouds, kPts);
  std::vector<float> pr(kClouds, 0.135f);
  for (int c = 0; c < kClouds; c++) pp[c] = pts.data() + (size_t)c * kPts * 3;
  std::vector<ssq_pairs> outs(kClouds);
  if (ssq_batch_reduce_points(h, kClouds, pp.data(), pn.data(), 3, pr.data(), 2, &cfg.points_params, outs.data(),
                              nullptr) == SSQ_OK)
    for (auto& o : outs) ssq_pairs_free(&o);
}

}  // namespace

struct ssq_server {
  ssq_server_config cfg;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue[3];  // [0] matrices, [1] point clouds, [2] registered matrices
  std::unordered_map<int64_t, Resident> resident;
  int64_t next_mid = 1;
  std::deque<Request*> sizing;    // point clouds whose cells are being counted
  std::condition_variable cv_size;
  std::vector<std::thread> sizers;
  std::unordered_map<int64_t, Result> done;
  std::vector<std::thread> workers;
  int64_t next_id = 1;
  int64_t in_flight_bytes = 0;
  int64_t outstanding = 0;
  int ready = 0;  // workers that finished their warm-up
  ssq_server_stats stats{};
  bool stop = false;
  int init_status = SSQ_OK;

  void worker_main(int wid) {
    ssq_handle* h = nullptr;
    int rc = ssq_create(cfg.device, &h);
    if (rc) {
      std::lock_guard<std::mutex> g(mu);
      init_status = rc;
    } else {
      ssq::handle_set_rips_streams(h, 1);
      warm_up(h, cfg);
    }
    {
      std::lock_guard<std::mutex> g(mu);
      ready++;
    }
    cv_done.notify_all();
    for (;;) {
      std::vector<Request*> batch;
      std::vector<ssq_device_matrix> dms;  // registered batches: the resident matrices
      int batch_kind = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        auto pending = [&] { return !queue[0].empty() || !queue[1].empty() || !queue[2].empty(); };
        cv_work.wait(lk, [&] { return (stop && sizing.empty()) || pending(); });
        if (!pending()) break;  // stopping, nothing left to size
        // one kind per batch (each has its own parameters and input path): the queue whose
        // head is oldest
        int kind = -1;
        for (int k = 0; k < 3; k++)
          if (!queue[k].empty() && (kind < 0 || queue[k].front()->t_submit < queue[kind].front()->t_submit)) kind = k;
        batch_kind = kind;
        auto& Q = queue[kind];
        // admission: oldest first while it fits the batch size and the device budget; the
        // first request is always admitted when nothing else is in flight
        int64_t cells = 0;
        while (!Q.empty() && (int)batch.size() < cfg.max_batch) {
          Request* q = Q.front();
          // point clouds of one batch share their dimension and simplex dimension
          if (kind == 1 && !batch.empty() && (q->pdim != batch[0]->pdim || q->maxdim != batch[0]->maxdim)) break;
          if (in_flight_bytes + q->est_bytes > cfg.mem_budget_bytes && (in_flight_bytes > 0 || !batch.empty()))
            break;
          if (cfg.max_batch_cells > 0 && !batch.empty() && cells + q->est_cells > cfg.max_batch_cells) break;
          cells += q->est_cells;
          Q.pop_front();
          in_flight_bytes += q->est_bytes;
          batch.push_back(q);
          if (kind == 2) dms.push_back(resident[q->mid].dm);  // held by the request's reference
        }
        if (batch.empty()) {
          // budget full: wait for another worker to return budget
          dms.clear();
          cv_done.wait(lk);
          continue;
        }
      }
      const auto t_start = Clock::now();
      int64_t batch_cells = 0;
      for (Request* q : batch) batch_cells += q->est_cells;
      std::vector<ssq_matrix> mats(batch.size());
      std::vector<int> st(batch.size(), SSQ_OK);
      const bool gpu_points = batch_kind == 1 && batch[0]->maxdim <= 2;
      const bool registered = batch_kind == 2;
      for (size_t i = 0; i < batch.size() && !gpu_points && !registered; i++) {
        Request* q = batch[i];
        if (q->is_points) {
          // dimension 3: the Rips matrix is built on the host
          st[i] = ssq::vr_build(q->pts.data(), q->npts, q->pdim, q->r, q->maxdim, &mats[i]);
        } else {
          mats[i].n = (int64_t)q->dims.size();
          mats[i].nnz = (int64_t)q->rows.size();
          mats[i].col_ptr = q->colptr.data();
          mats[i].rows = q->rows.data();
          mats[i].dims = q->dims.data();
        }
      }
      std::vector<const ssq_matrix*> ptrs;
      std::vector<size_t> idx;
      for (size_t i = 0; i < batch.size(); i++)
        if (st[i] == SSQ_OK) {
          ptrs.push_back(&mats[i]);
          idx.push_back(i);
        }
      std::vector<ssq_pairs> outs(ptrs.size());
      int brc = SSQ_OK;
      if (gpu_points) {
        // the Rips matrices are built on this worker's stream straight into one batch in HBM
        std::vector<const float*> pp(batch.size());
        std::vector<int64_t> pn(batch.size());
        std::vector<float> pr(batch.size());
        for (size_t i = 0; i < batch.size(); i++) {
          pp[i] = batch[i]->pts.data();
          pn[i] = batch[i]->npts;
          pr[i] = batch[i]->r;
        }
        outs.resize(batch.size());
        brc = h ? ssq_batch_reduce_points(h, (int)batch.size(), pp.data(), pn.data(), batch[0]->pdim, pr.data(),
                                          batch[0]->maxdim, &cfg.points_params, outs.data(), nullptr)
                : SSQ_EDEVICE;
      } else if (registered) {
        // gathered from their resident copies on this worker's stream
        brc = h ? ssq_batch_reduce_resident(h, dms.data(), (int)dms.size(), &cfg.params, outs.data(), nullptr)
                : SSQ_EDEVICE;
      } else if (!ptrs.empty()) {
        const ssq_params* bp = batch_kind ? &cfg.points_params : &cfg.params;
        brc = h ? ssq_batch_reduce(h, ptrs.data(), (int)ptrs.size(), bp, outs.data(), nullptr) : SSQ_EDEVICE;
      }
      if (brc && h && batch.size() > 1) {
        // a failed batch is reduced again request by request, so that only a bad request
        // reports the error
        brc = SSQ_OK;
        outs.assign(gpu_points ? batch.size() : ptrs.size(), ssq_pairs{});
        for (size_t k = 0; k < outs.size(); k++) {
          const size_t i = gpu_points ? k : idx[k];
          Request* q = batch[i];
          int rc1;
          if (gpu_points) {
            const float* p1 = q->pts.data();
            rc1 = ssq_batch_reduce_points(h, 1, &p1, &q->npts, q->pdim, &q->r, q->maxdim, &cfg.points_params,
                                          &outs[k], nullptr);
          } else if (registered) {
            rc1 = ssq_batch_reduce_resident(h, &dms[i], 1, &cfg.params, &outs[k], nullptr);
          } else {
            const ssq_params* bp = batch_kind ? &cfg.points_params : &cfg.params;
            rc1 = ssq_batch_reduce(h, &ptrs[k], 1, bp, &outs[k], nullptr);
          }
          if (rc1) {
            st[i] = rc1;
            outs[k] = ssq_pairs{};
          }
        }
        if (gpu_points) {  // results of the retried clouds, by request
          idx.clear();
          for (size_t i = 0; i < batch.size(); i++) idx.push_back(i);
        }
      }
      const auto t_end = Clock::now();
      {
        std::lock_guard<std::mutex> g(mu);
        for (size_t i = 0; i < batch.size(); i++) {
          Result& R = done[batch[i]->id];
          R.latency_ms = std::chrono::duration<double, std::milli>(t_end - batch[i]->t_submit).count();
          R.status = st[i];
        }
        for (size_t k = 0; k < idx.size(); k++) {
          Result& R = done[batch[idx[k]]->id];
          if (brc) {
            R.status = brc;
            continue;
          }
          if (R.status) continue;  // failed alone
          R.pairs.assign(outs[k].pairs, outs[k].pairs + 2 * outs[k].np);
          R.ess.assign(outs[k].essential, outs[k].essential + outs[k].ne);
          ssq_pairs_free(&outs[k]);
        }
        const double service = std::chrono::duration<double, std::milli>(t_end - t_start).count();
        stats.batches[batch_kind]++;
        stats.requests[batch_kind] += (int64_t)batch.size();
        stats.cells[batch_kind] += batch_cells;
        stats.service_ms[batch_kind] += service;
        stats.max_service_ms[batch_kind] = std::max(stats.max_service_ms[batch_kind], service);
        for (Request* q : batch) {
          in_flight_bytes -= q->est_bytes;
          outstanding--;
          if (q->mid) resident[q->mid].refs--;
        }
      }
      for (size_t i = 0; i < batch.size() && !gpu_points && !registered; i++)
        if (batch[i]->is_points && st[i] == SSQ_OK) ssq_matrix_free(&mats[i]);
      for (Request* q : batch) delete q;
      cv_done.notify_all();
      cv_work.notify_all();
    }
    if (h) ssq_destroy(h);
    (void)wid;
  }

  // Point clouds are counted (edges, triangles) by sizer threads before they are queued, so
  // that a large cloud's count never holds up the submitting thread; the latency clock starts
  // at submission.
  int submit(Request* q, int64_t* ticket) {
    {
      std::lock_guard<std::mutex> g(mu);
      if (stop) {
        delete q;
        return SSQ_EINVAL;
      }
      q->id = next_id++;
      q->t_submit = Clock::now();
      if (q->mid) {
        auto it = resident.find(q->mid);
        if (it == resident.end() || it->second.closing) {
          delete q;
          return SSQ_EINVAL;
        }
        it->second.refs++;
        q->est_bytes = it->second.est_bytes;
        q->est_cells = it->second.dm.n;
      }
      (q->is_points ? sizing : q->mid ? queue[2] : queue[0]).push_back(q);
      outstanding++;
      *ticket = q->id;
    }
    (q->is_points ? cv_size : cv_work).notify_one();
    return SSQ_OK;
  }

  void sizer_main() {
    for (;;) {
      Request* q;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_size.wait(lk, [&] { return stop || !sizing.empty(); });
        if (sizing.empty()) break;  // stopping
        q = sizing.front();
        sizing.pop_front();
      }
      estimate_points(q->pts.data(), q->npts, q->pdim, q->r, q->maxdim, &q->est_cells, &q->est_bytes);
      {
        std::lock_guard<std::mutex> g(mu);
        queue[1].push_back(q);
      }
      cv_work.notify_one();
    }
  }
};

extern "C" int ssq_server_create(const ssq_server_config* cfg, ssq_server** out) {
  if (!cfg || !out || cfg->streams < 1 || cfg->max_batch < 1) return SSQ_EINVAL;
  *out = nullptr;
  ssq_server* s = new (std::nothrow) ssq_server();
  if (!s) return SSQ_ENOMEM;
  s->cfg = *cfg;
  if (s->cfg.mem_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipSetDevice(cfg->device) != hipSuccess || hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return SSQ_EDEVICE;
    }
    s->cfg.mem_budget_bytes = (int64_t)(fr / 2);
  }
  for (int w = 0; w < cfg->streams; w++) s->workers.emplace_back(&ssq_server::worker_main, s, w);
  const int nsizers = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
  for (int k = 0; k < nsizers; k++) s->sizers.emplace_back(&ssq_server::sizer_main, s);
  {  // requests are accepted once every worker has sized its buffers
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->ready == cfg->streams; });
  }
  *out = s;
  return SSQ_OK;
}

extern "C" int ssq_server_submit_matrix(ssq_server* s, const ssq_matrix* m, int64_t* ticket) {
  if (!s || !m || !ticket) return SSQ_EINVAL;
  Request* q = new (std::nothrow) Request();
  if (!q) return SSQ_ENOMEM;
  if (!matrix_ok(m)) {
    delete q;
    return SSQ_EFORMAT;
  }
  q->colptr.assign(m->col_ptr, m->col_ptr + m->n + 1);
  if (m->nnz) q->rows.assign(m->rows, m->rows + m->nnz);
  q->dims.assign(m->dims, m->dims + m->n);
  q->est_bytes = estimate_bytes(m->n, m->nnz);
  q->est_cells = m->n;
  return s->submit(q, ticket);
}

extern "C" int ssq_server_submit_points(ssq_server* s, const float* pts, int64_t npts, int pdim, float r, int maxdim,
                                        int64_t* ticket) {
  if (!s || !ticket || npts < 0 || pdim < 1 || (npts && !pts) || maxdim < 0 || maxdim > 3) return SSQ_EINVAL;
  Request* q = new (std::nothrow) Request();
  if (!q) return SSQ_ENOMEM;
  q->is_points = true;
  q->pts.assign(pts, pts + npts * pdim);
  q->npts = npts;
  q->pdim = pdim;
  q->r = r;
  q->maxdim = maxdim;
  return s->submit(q, ticket);
}

extern "C" int ssq_server_register_matrix(ssq_server* s, const ssq_matrix* m, int64_t* matrix_id) {
  if (!s || !m || !matrix_id) return SSQ_EINVAL;
  *matrix_id = 0;
  if (!matrix_ok(m) || !rows_ok(m)) return SSQ_EFORMAT;
  if (hipSetDevice(s->cfg.device) != hipSuccess) return SSQ_EDEVICE;
  // synchronous copies on the null stream; the workers' streams do not wait for it
  Resident R;
  void *cp = nullptr, *rw = nullptr, *dm = nullptr;
  const size_t bc = sizeof(int64_t) * (size_t)(m->n + 1), br = sizeof(int32_t) * (size_t)m->nnz,
               bd = sizeof(int32_t) * (size_t)m->n;
  bool ok = hipMalloc(&cp, bc) == hipSuccess && hipMalloc(&rw, std::max<size_t>(br, 4)) == hipSuccess &&
            hipMalloc(&dm, std::max<size_t>(bd, 4)) == hipSuccess;
  if (ok && m->n == 0) {
    const int64_t zero = 0;
    ok = hipMemcpy(cp, &zero, sizeof(zero), hipMemcpyHostToDevice) == hipSuccess;
  } else if (ok) {
    ok = hipMemcpy(cp, m->col_ptr, bc, hipMemcpyHostToDevice) == hipSuccess &&
         (br == 0 || hipMemcpy(rw, m->rows, br, hipMemcpyHostToDevice) == hipSuccess) &&
         hipMemcpy(dm, m->dims, bd, hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    for (void* p : {cp, rw, dm})
      if (p) (void)hipFree(p);
    return SSQ_ENOMEM;
  }
  R.dm = ssq_device_matrix{m->n, m->nnz, (const int64_t*)cp, (const int32_t*)rw, (const int32_t*)dm};
  R.est_bytes = estimate_bytes(m->n, m->nnz);
  std::lock_guard<std::mutex> g(s->mu);
  *matrix_id = s->next_mid++;
  s->resident[*matrix_id] = R;
  return SSQ_OK;
}

extern "C" int ssq_server_submit_registered(ssq_server* s, int64_t matrix_id, int64_t* ticket) {
  if (!s || !ticket || matrix_id <= 0) return SSQ_EINVAL;
  Request* q = new (std::nothrow) Request();
  if (!q) return SSQ_ENOMEM;
  q->mid = matrix_id;
  return s->submit(q, ticket);
}

extern "C" int ssq_server_unregister_matrix(ssq_server* s, int64_t matrix_id) {
  if (!s) return SSQ_EINVAL;
  ssq_device_matrix dm;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->resident.find(matrix_id);
    if (it == s->resident.end() || it->second.closing) return SSQ_EINVAL;
    it->second.closing = true;  // no new requests on it
    s->cv_done.wait(lk, [&] { return s->resident[matrix_id].refs == 0; });
    dm = s->resident[matrix_id].dm;
    s->resident.erase(matrix_id);
  }
  if (hipSetDevice(s->cfg.device) != hipSuccess) return SSQ_EDEVICE;
  for (const void* p : {(const void*)dm.col_ptr, (const void*)dm.rows, (const void*)dm.dims})
    (void)hipFree(const_cast<void*>(p));
  return SSQ_OK;
}

extern "C" int ssq_server_poll(ssq_server* s, int64_t ticket, ssq_pairs* out, double* latency_ms) {
  if (!s || !out) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->done.find(ticket);
  if (it == s->done.end()) return SSQ_EBUSY;
  Result R = std::move(it->second);
  s->done.erase(it);
  if (latency_ms) *latency_ms = R.latency_ms;
  if (R.status) return R.status;
  out->np = (int64_t)R.pairs.size() / 2;
  out->ne = (int64_t)R.ess.size();
  out->pairs = (int64_t*)std::malloc(sizeof(int64_t) * std::max<size_t>(1, R.pairs.size()));
  out->essential = (int64_t*)std::malloc(sizeof(int64_t) * std::max<size_t>(1, R.ess.size()));
  if (!out->pairs || !out->essential) {
    ssq_pairs_free(out);
    return SSQ_ENOMEM;
  }
  if (!R.pairs.empty()) std::memcpy(out->pairs, R.pairs.data(), sizeof(int64_t) * R.pairs.size());
  if (!R.ess.empty()) std::memcpy(out->essential, R.ess.data(), sizeof(int64_t) * R.ess.size());
  return SSQ_OK;
}

extern "C" int ssq_server_get_stats(ssq_server* s, ssq_server_stats* out) {
  if (!s || !out) return SSQ_EINVAL;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  return SSQ_OK;
}

extern "C" int ssq_server_wait_all(ssq_server* s) {
  if (!s) return SSQ_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
  return s->init_status;
}

extern "C" int ssq_server_destroy(ssq_server* s) {
  if (!s) return SSQ_EINVAL;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_size.notify_all();
  for (auto& t : s->sizers) t.join();  // they drain the sizing queue first
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& Q : s->queue)
    for (Request* q : Q) delete q;
  if (!s->resident.empty() && hipSetDevice(s->cfg.device) == hipSuccess)
    for (auto& kv : s->resident)
      for (const void* p : {(const void*)kv.second.dm.col_ptr, (const void*)kv.second.dm.rows,
                            (const void*)kv.second.dm.dims})
        (void)hipFree(const_cast<void*>(p));
  delete s;
  return SSQ_OK;
}
