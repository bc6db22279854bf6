// PHAT boundary-matrix reader and persistence-pair writer (host side of libssq).
//
// Replaces processFiles (main.cu:90-167) and writeVectorToFile (main.cu:191-229).
// Same accepted syntax: ASCII lines "dim b0 b1 ...", skipping lines that are empty after
// right-trim or start with '#'; binary int64 stream n, {dim, k, idx[k]}.  Boundaries are
// sorted and de-duplicated on read (main.cu:118-119,149-150).  Unlike the reference, both
// formats keep the file's 0-based indices (no +1 shift, no dummy column) and every index
// is range-checked.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "ssq_internal.h"

namespace {

struct Builder {
  std::vector<int64_t> colptr{0};
  std::vector<int32_t> rows;
  std::vector<int32_t> dims;
  void push(int32_t dim, int32_t* r, int64_t k) {
    std::sort(r, r + k);
    int64_t u = std::unique(r, r + k) - r;
    rows.insert(rows.end(), r, r + u);
    colptr.push_back((int64_t)rows.size());
    dims.push_back(dim);
  }
};

int finish(Builder& b, ssq_matrix* out) {
  int64_t n = (int64_t)b.dims.size();
  for (int32_t r : b.rows)
    if (r < 0 || r >= n) return SSQ_EFORMAT;
  return ssq::matrix_from_vectors(b.colptr, b.rows, b.dims, out);
}

int read_all(const char* path, std::vector<char>& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return SSQ_EIO;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0) {
    std::fclose(f);
    return SSQ_EIO;
  }
  buf.resize((size_t)sz + 1);
  size_t got = sz ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
  std::fclose(f);
  if (got != (size_t)sz) return SSQ_EIO;
  buf[(size_t)sz] = '\n';
  return SSQ_OK;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v'; }

}  // namespace

namespace ssq {

int matrix_from_vectors(const std::vector<int64_t>& colptr, const std::vector<int32_t>& rows,
                        const std::vector<int32_t>& dims, ssq_matrix* out) {
  int64_t n = (int64_t)dims.size();
  out->n = n;
  out->nnz = (int64_t)rows.size();
  out->col_ptr = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  out->rows = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, out->nnz));
  out->dims = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, n));
  if (!out->col_ptr || !out->rows || !out->dims) {
    ssq_matrix_free(out);
    return SSQ_ENOMEM;
  }
  std::memcpy(out->col_ptr, colptr.data(), sizeof(int64_t) * (size_t)(n + 1));
  if (out->nnz) std::memcpy(out->rows, rows.data(), sizeof(int32_t) * (size_t)out->nnz);
  if (n) std::memcpy(out->dims, dims.data(), sizeof(int32_t) * (size_t)n);
  return SSQ_OK;
}

}  // namespace ssq

extern "C" int ssq_load_phat(const char* path, int binary, ssq_matrix* out) {
  if (!path || !out) return SSQ_EINVAL;
  std::memset(out, 0, sizeof(*out));
  std::vector<char> buf;
  int rc = read_all(path, buf);
  if (rc) return rc;
  Builder b;
  std::vector<int32_t> tmp;
  if (!binary) {
    const char* p = buf.data();
    const char* end = buf.data() + buf.size();
    while (p < end) {
      const char* eol = (const char*)std::memchr(p, '\n', (size_t)(end - p));
      if (!eol) eol = end;
      const char* q = eol;
      while (q > p && (is_space(q[-1]) || q[-1] == '\n')) q--;
      if (q > p && p[0] != '#') {
        // parse integers in [p, q)
        tmp.clear();
        const char* s = p;
        bool have_dim = false;
        int32_t dim = 0;
        while (s < q) {
          while (s < q && is_space(*s)) s++;
          if (s >= q) break;
          bool neg = false;
          if (*s == '-' || *s == '+') {
            neg = (*s == '-');
            s++;
          }
          if (s >= q || *s < '0' || *s > '9') return SSQ_EFORMAT;
          int64_t v = 0;
          while (s < q && *s >= '0' && *s <= '9') {
            v = v * 10 + (*s - '0');
            if (v > 0x7fffffffLL) return SSQ_EFORMAT;
            s++;
          }
          if (s < q && !is_space(*s)) return SSQ_EFORMAT;
          if (neg) v = -v;
          if (!have_dim) {
            dim = (int32_t)v;
            have_dim = true;
          } else {
            if (v < 0) return SSQ_EFORMAT;
            tmp.push_back((int32_t)v);
          }
        }
        if (!have_dim || dim < 0) return SSQ_EFORMAT;
        b.push(dim, tmp.data(), (int64_t)tmp.size());
      }
      p = eol + 1;
    }
  } else {
    size_t sz = buf.size() - 1;
    const char* p = buf.data();
    size_t pos = 0;
    auto rd = [&](int64_t& v) -> bool {
      if (pos + 8 > sz) return false;
      std::memcpy(&v, p + pos, 8);
      pos += 8;
      return true;
    };
    int64_t nc;
    if (!rd(nc) || nc < 0) return SSQ_EFORMAT;
    for (int64_t c = 0; c < nc; c++) {
      int64_t dim, k;
      if (!rd(dim) || !rd(k) || k < 0 || dim < 0 || dim > 0x7fffffffLL) return SSQ_EFORMAT;
      if ((size_t)k > (sz - pos) / 8) return SSQ_EFORMAT;
      tmp.resize((size_t)k);
      for (int64_t i = 0; i < k; i++) {
        int64_t v;
        rd(v);
        if (v < 0 || v > 0x7fffffffLL) return SSQ_EFORMAT;
        tmp[(size_t)i] = (int32_t)v;
      }
      b.push((int32_t)dim, tmp.data(), k);
    }
  }
  return finish(b, out);
}

extern "C" int ssq_save_pairs(const char* path, int binary, const ssq_pairs* pr) {
  if (!path || !pr) return SSQ_EINVAL;
  FILE* f = std::fopen(path, binary ? "wb" : "w");
  if (!f) return SSQ_EIO;
  bool ok = true;
  if (binary) {
    ok = std::fwrite(&pr->np, 8, 1, f) == 1;
    if (ok && pr->np) ok = std::fwrite(pr->pairs, 8, (size_t)(2 * pr->np), f) == (size_t)(2 * pr->np);
  } else {
    std::vector<char> out;
    out.reserve((size_t)(pr->np * 16 + 32));
    char tmp[48];
    int l = std::snprintf(tmp, sizeof(tmp), "%lld\n", (long long)pr->np);
    out.insert(out.end(), tmp, tmp + l);
    for (int64_t i = 0; i < pr->np; i++) {
      l = std::snprintf(tmp, sizeof(tmp), "%lld %lld\n", (long long)pr->pairs[2 * i],
                        (long long)pr->pairs[2 * i + 1]);
      out.insert(out.end(), tmp, tmp + l);
    }
    ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? SSQ_OK : SSQ_EIO;
}

extern "C" void ssq_matrix_free(ssq_matrix* m) {
  if (!m) return;
  std::free(m->col_ptr);
  std::free(m->rows);
  std::free(m->dims);
  std::memset(m, 0, sizeof(*m));
}

extern "C" void ssq_pairs_free(ssq_pairs* p) {
  if (!p) return;
  std::free(p->pairs);
  std::free(p->essential);
  std::memset(p, 0, sizeof(*p));
}

extern "C" void ssq_default_params(ssq_params* p) {
  p->dualize = 0;
  p->model = 0;
  p->max_iteration = 10;
  p->ss_block_size = 102400;
  p->growth = 1.05f;
  p->queue_size = 65536;
  p->level1 = 1000;
  p->level2 = 2000;
  p->add_waves = 0;
  p->long_grid = 0;
  p->short_max = 0;
  p->rounds_per_sync = 0;
}
