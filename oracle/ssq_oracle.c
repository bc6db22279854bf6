/*
 * ssq oracle — CPU restatement of the SpecSeq++ reduction path (checker only).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg. The product library never links or calls anything in this file.
 * Function-by-function provenance and pinning are listed in ssq_oracle.h.
 */
#include "ssq_oracle.h"

#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------------------ */
/* small helpers                                                                         */
/* ------------------------------------------------------------------------------------ */
static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}
static int64_t sort_unique_i32(int32_t* v, int64_t k) {
  if (k <= 1) return k;
  qsort(v, (size_t)k, sizeof(int32_t), cmp_i32);
  int64_t o = 1;
  for (int64_t i = 1; i < k; i++)
    if (v[i] != v[o - 1]) v[o++] = v[i];
  return o;
}

typedef struct {
  int64_t n, cap_cols, nnz, cap_nnz;
  int64_t* colptr;
  int32_t* rows;
  int32_t* dims;
} mbuild;

static int mb_push_col(mbuild* m, int32_t dim, const int32_t* r, int64_t k) {
  if (m->n + 2 > m->cap_cols) {
    m->cap_cols = m->cap_cols ? m->cap_cols * 2 : 1024;
    m->colptr = (int64_t*)realloc(m->colptr, sizeof(int64_t) * (size_t)m->cap_cols);
    m->dims = (int32_t*)realloc(m->dims, sizeof(int32_t) * (size_t)m->cap_cols);
    if (!m->colptr || !m->dims) return -1;
  }
  if (m->nnz + k > m->cap_nnz) {
    while (m->nnz + k > m->cap_nnz) m->cap_nnz = m->cap_nnz ? m->cap_nnz * 2 : 4096;
    m->rows = (int32_t*)realloc(m->rows, sizeof(int32_t) * (size_t)m->cap_nnz);
    if (!m->rows) return -1;
  }
  if (m->n == 0) m->colptr[0] = 0;
  if (k > 0) memcpy(m->rows + m->nnz, r, sizeof(int32_t) * (size_t)k);
  int64_t kk = sort_unique_i32(m->rows + m->nnz, k); /* main.cu:118-119 sort + unique */
  m->nnz += kk;
  m->dims[m->n] = dim;
  m->n++;
  m->colptr[m->n] = m->nnz;
  return 0;
}

/* ------------------------------------------------------------------------------------ */
/* PHAT I/O (main.cu:90-167, 191-229)                                                    */
/* ------------------------------------------------------------------------------------ */
int or_load_phat(const char* path, int binary, int64_t* n_out, int64_t** colptr_out,
                 int32_t** rows_out, int32_t** dims_out) {
  FILE* f = fopen(path, binary ? "rb" : "r");
  if (!f) return 1;
  mbuild m;
  memset(&m, 0, sizeof(m));
  int rc = 0;
  int32_t* tmp = NULL;
  int64_t tcap = 0;
  if (!binary) {
    /* One cell per line: dim then boundary indices; lines empty after right-trim or whose
     * first character is '#' are skipped (main.cu:104-106). */
    size_t lcap = 1 << 16;
    char* line = (char*)malloc(lcap);
    for (;;) {
      size_t len = 0;
      int ch;
      while ((ch = fgetc(f)) != EOF && ch != '\n') {
        if (len + 1 >= lcap) {
          lcap *= 2;
          line = (char*)realloc(line, lcap);
        }
        line[len++] = (char)ch;
      }
      if (ch == EOF && len == 0) break;
      line[len] = 0;
      while (len > 0 && isspace((unsigned char)line[len - 1])) line[--len] = 0;
      if (len == 0 || line[0] == '#') {
        if (ch == EOF) break;
        continue;
      }
      char* p = line;
      char* e;
      long dim = strtol(p, &e, 10);
      if (e == p) { rc = 2; break; }
      p = e;
      int64_t k = 0;
      for (;;) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        long long v = strtoll(p, &e, 10);
        if (e == p) { rc = 2; break; }
        p = e;
        if (k + 1 > tcap) {
          tcap = tcap ? tcap * 2 : 64;
          tmp = (int32_t*)realloc(tmp, sizeof(int32_t) * (size_t)tcap);
        }
        if (v < 0 || v > 0x7fffffffLL) { rc = 2; break; }
        tmp[k++] = (int32_t)v;
      }
      if (rc) break;
      if (mb_push_col(&m, (int32_t)dim, tmp, k)) { rc = 2; break; }
      if (ch == EOF) break;
    }
    free(line);
  } else {
    /* int64 ncols, then per column int64 dim, int64 k, int64 rows[k] (main.cu:135-160). */
    int64_t nc;
    if (fread(&nc, 8, 1, f) != 1 || nc < 0) rc = 2;
    for (int64_t c = 0; !rc && c < nc; c++) {
      int64_t hd[2];
      if (fread(hd, 8, 2, f) != 2 || hd[1] < 0) { rc = 2; break; }
      int64_t k = hd[1];
      if (k > tcap) {
        tcap = k;
        tmp = (int32_t*)realloc(tmp, sizeof(int32_t) * (size_t)(tcap ? tcap : 1));
      }
      for (int64_t i = 0; i < k; i++) {
        int64_t v;
        if (fread(&v, 8, 1, f) != 1 || v < 0 || v > 0x7fffffffLL) { rc = 2; break; }
        tmp[i] = (int32_t)v;
      }
      if (rc) break;
      if (mb_push_col(&m, (int32_t)hd[0], tmp, k)) rc = 2;
    }
  }
  fclose(f);
  free(tmp);
  if (!rc) {
    if (m.n == 0) {
      m.colptr = (int64_t*)calloc(1, sizeof(int64_t));
      m.dims = (int32_t*)calloc(1, sizeof(int32_t));
    }
    for (int64_t i = 0; i < m.nnz; i++)
      if (m.rows[i] >= m.n) { rc = 2; break; }
  }
  if (rc) {
    free(m.colptr);
    free(m.rows);
    free(m.dims);
    return rc;
  }
  if (!m.rows) m.rows = (int32_t*)calloc(1, sizeof(int32_t));
  *n_out = m.n;
  *colptr_out = m.colptr;
  *rows_out = m.rows;
  *dims_out = m.dims;
  return 0;
}

void or_free(void* p) { free(p); }

int or_save_pairs(const char* path, int binary, int64_t np, const int64_t* pairs) {
  FILE* f = fopen(path, binary ? "wb" : "w");
  if (!f) return 1;
  if (binary) {
    fwrite(&np, 8, 1, f);
    fwrite(pairs, 8, (size_t)(2 * np), f);
  } else {
    fprintf(f, "%lld\n", (long long)np);
    for (int64_t i = 0; i < np; i++)
      fprintf(f, "%lld %lld\n", (long long)pairs[2 * i], (long long)pairs[2 * i + 1]);
  }
  fclose(f);
  return 0;
}

/* Pairs sorted by birth: births are distinct, so a bucket pass over [0,n) sorts them
 * (specseq.h:1182-1193 sorts (birth, death) pairs). */
static int64_t emit_pairs(int64_t n, const int32_t* death_of_birth, int64_t* out) {
  int64_t np = 0;
  for (int64_t b = 0; b < n; b++)
    if (death_of_birth[b] >= 0) {
      out[2 * np] = b;
      out[2 * np + 1] = death_of_birth[b];
      np++;
    }
  return np;
}

/* ------------------------------------------------------------------------------------ */
/* Bit-tree pivot column (PHAT's default representation, phat.cpp:184): a 64-ary tree of   */
/* bit sets with O(depth) toggle and O(depth) maximum.                                    */
/* ------------------------------------------------------------------------------------ */
typedef struct {
  int nlev;
  int64_t nw[8];
  uint64_t* w[8];
  int64_t count;
} bitree;

static void bt_init(bitree* t, int64_t nbits) {
  memset(t, 0, sizeof(*t));
  int64_t words = (nbits + 63) / 64;
  if (words < 1) words = 1;
  for (;;) {
    t->nw[t->nlev] = words;
    t->w[t->nlev] = (uint64_t*)calloc((size_t)words, 8);
    t->nlev++;
    if (words == 1) break;
    words = (words + 63) / 64;
  }
}
static void bt_free(bitree* t) {
  for (int l = 0; l < t->nlev; l++) free(t->w[l]);
}
static inline void bt_toggle(bitree* t, int64_t i) {
  uint64_t* w0 = &t->w[0][i >> 6];
  uint64_t before = *w0;
  *w0 ^= 1ULL << (i & 63);
  t->count += (*w0 >> (i & 63)) & 1 ? 1 : -1;
  int64_t idx = i >> 6;
  for (int l = 1; l < t->nlev; l++) {
    uint64_t now = t->w[l - 1][idx];
    if ((before != 0) == (now != 0)) break;
    uint64_t* wl = &t->w[l][idx >> 6];
    before = *wl;
    *wl ^= 1ULL << (idx & 63);
    idx >>= 6;
  }
}
static inline int64_t bt_max(const bitree* t) {
  int l = t->nlev - 1;
  if (!t->w[l][0]) return -1;
  int64_t idx = 0;
  for (; l >= 0; l--) {
    uint64_t word = t->w[l][idx];
    idx = idx * 64 + (63 - __builtin_clzll(word));
  }
  return idx;
}

/* ------------------------------------------------------------------------------------ */
/* Twist reduction (phat.cpp:184-186 default algorithm; Chen-Kerber twist): dimensions     */
/* from high to low; after column j gets pivot low, column low is cleared.                */
/* ------------------------------------------------------------------------------------ */
int64_t or_twist_reduce(int64_t n, const int64_t* colptr, const int32_t* rows,
                        const int32_t* dims, const uint8_t* skip, int64_t* pairs_out,
                        double* stats_out) {
  int32_t maxd = 0;
  if (n < 0) return -1;
  for (int64_t c = 0; c < n; c++)
    if (dims[c] > maxd) maxd = dims[c];
  int32_t* pivot_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1)); /* row -> col */
  uint8_t* cleared = (uint8_t*)calloc((size_t)(n ? n : 1), 1);
  int64_t* roff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  int32_t* rlen = (int32_t*)calloc((size_t)(n ? n : 1), sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) pivot_of[i] = -1;
  int64_t acap = 1 << 20, aused = 0;
  int32_t* arena = (int32_t*)malloc(sizeof(int32_t) * (size_t)acap);
  int32_t* ext = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  double adds = 0, bytes = 0, max_chain = 0, critical = 0;
  /* fin[c]: when column c would be final if every addition took one step and a column could
   * add a pivot only once that pivot is final (unlimited parallelism, dimensions in turn) */
  double* fin = (double*)calloc((size_t)(n ? n : 1), sizeof(double));
  bitree t;
  bt_init(&t, n);
  for (int32_t d = maxd; d >= 1; d--) {
    double dim_critical = 0;
    for (int64_t j = 0; j < n; j++) {
      if (dims[j] != d || cleared[j] || (skip && skip[j])) continue;
      for (int64_t k = colptr[j]; k < colptr[j + 1]; k++) bt_toggle(&t, rows[k]);
      int64_t low;
      double col_adds = 0, tj = 0;
      while ((low = bt_max(&t)) >= 0 && pivot_of[low] >= 0) {
        col_adds += 1;
        int32_t p = pivot_of[low];
        tj = (fin[p] > tj ? fin[p] : tj) + 1;
        int64_t a = t.count;
        const int32_t* col = arena + roff[p];
        for (int32_t k = 0; k < rlen[p]; k++) bt_toggle(&t, col[k]);
        adds += 1;
        bytes += 4.0 * (double)(a + rlen[p] + t.count) + 24.0;
      }
      if (col_adds > max_chain) max_chain = col_adds;
      fin[j] = tj;
      if (tj > dim_critical) dim_critical = tj;
      if (low < 0) continue;
      /* store the reduced column (descending extraction, then reverse) */
      int64_t len = 0;
      int64_t m;
      while ((m = bt_max(&t)) >= 0) {
        ext[len++] = (int32_t)m;
        bt_toggle(&t, m);
      }
      if (aused + len > acap) {
        while (aused + len > acap) acap *= 2;
        arena = (int32_t*)realloc(arena, sizeof(int32_t) * (size_t)acap);
      }
      for (int64_t k = 0; k < len; k++) arena[aused + k] = ext[len - 1 - k];
      roff[j] = aused;
      rlen[j] = (int32_t)len;
      aused += len;
      pivot_of[low] = (int32_t)j;
      cleared[low] = 1; /* twist clearing: column low is positive, reduces to zero */
    }
    critical += dim_critical;
  }
  bt_free(&t);
  free(fin);
  int32_t* dob = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; i++) dob[i] = pivot_of[i];
  int64_t np = emit_pairs(n, dob, pairs_out);
  if (stats_out) {
    stats_out[0] = adds;
    stats_out[1] = bytes;
    stats_out[2] = max_chain;
    stats_out[3] = critical;
  }
  free(dob);
  free(pivot_of);
  free(cleared);
  free(roff);
  free(rlen);
  free(arena);
  free(ext);
  return np;
}

/* ------------------------------------------------------------------------------------ */
/* Standard algorithm on dense bit columns: left to right, no clearing, no dim order.      */
/* Cohomology: the same twist reduction on the anti-transpose, pairs mapped back
 * (specseq.h:303-363 builds the anti-transpose when dualize is set; specseq.h:1188 maps a dual
 * pair (b', d') to (n-1-d', n-1-b')).  Dual cell n-1-c has the coboundary of c as its column,
 * rows ascending, and dimension maxdim - dim(c).  Skipped cells and their entries are left out
 * of the dual. */
int64_t or_twist_reduce_dual(int64_t n, const int64_t* colptr, const int32_t* rows,
                             const int32_t* dims, const uint8_t* skip, int64_t* pairs_out,
                             double* stats_out) {
  if (n < 0) return -1;
  int32_t maxd = 0;
  for (int64_t c = 0; c < n; c++)
    if (dims[c] > maxd) maxd = dims[c];
  const int64_t nnz = colptr[n];
  int64_t* dcp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int32_t* drw = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  int32_t* ddm = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  uint8_t* dsk = skip ? (uint8_t*)malloc((size_t)(n ? n : 1)) : NULL;
  int64_t* dob = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)(n ? n : 1));
  int32_t* death = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  /* counts: entry (r, c) of the primal is entry (n-1-c, n-1-r) of the dual */
  for (int64_t c = 0; c < n; c++) {
    if (skip && skip[c]) continue;
    for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) dcp[n - 1 - rows[k] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) dcp[i + 1] += dcp[i];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; i++) fill[i] = dcp[i];
  /* primal columns from the last to the first: each dual column's rows come out ascending */
  for (int64_t c = n - 1; c >= 0; c--) {
    if (skip && skip[c]) continue;
    for (int64_t k = colptr[c]; k < colptr[c + 1]; k++) {
      const int64_t dc = n - 1 - rows[k];
      drw[fill[dc]++] = (int32_t)(n - 1 - c);
    }
  }
  for (int64_t c = 0; c < n; c++) {
    ddm[n - 1 - c] = maxd - dims[c];
    if (dsk) dsk[n - 1 - c] = skip[c];
  }
  const int64_t np = or_twist_reduce(n, dcp, drw, ddm, dsk, dob, stats_out);
  for (int64_t i = 0; i < n; i++) death[i] = -1;
  for (int64_t k = 0; k < np; k++) death[n - 1 - dob[2 * k + 1]] = (int32_t)(n - 1 - dob[2 * k]);
  const int64_t np2 = emit_pairs(n, death, pairs_out);
  free(dcp);
  free(drw);
  free(ddm);
  free(dsk);
  free(dob);
  free(death);
  free(fill);
  return np2;
}

/* ------------------------------------------------------------------------------------ */
int64_t or_dense_standard(int64_t n, const int64_t* colptr, const int32_t* rows,
                          const uint8_t* skip, int64_t* pairs_out) {
  if (n > 65536 || n < 0) return -1;
  int64_t W = (n + 63) / 64;
  if (W < 1) W = 1;
  uint64_t* M = (uint64_t*)calloc((size_t)(n ? n : 1) * (size_t)W, 8);
  int32_t* low_owner = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  int32_t* dob = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; i++) low_owner[i] = dob[i] = -1;
  for (int64_t j = 0; j < n; j++) {
    uint64_t* cj = M + j * W;
    if (!(skip && skip[j]))
      for (int64_t k = colptr[j]; k < colptr[j + 1]; k++) cj[rows[k] >> 6] ^= 1ULL << (rows[k] & 63);
    for (;;) {
      int64_t low = -1;
      for (int64_t w = W - 1; w >= 0; w--)
        if (cj[w]) {
          low = w * 64 + 63 - __builtin_clzll(cj[w]);
          break;
        }
      if (low < 0) break;
      int32_t p = low_owner[low];
      if (p < 0) {
        low_owner[low] = (int32_t)j;
        dob[low] = (int32_t)j;
        break;
      }
      const uint64_t* cp = M + (int64_t)p * W;
      for (int64_t w = 0; w < W; w++) cj[w] ^= cp[w];
    }
  }
  int64_t np = emit_pairs(n, dob, pairs_out);
  free(M);
  free(low_owner);
  free(dob);
  return np;
}

/* ------------------------------------------------------------------------------------ */
/* Rank formula: pairs from GF(2) ranks of lower-left submatrices (no reduction).          */
/* ------------------------------------------------------------------------------------ */
int64_t or_rank_pairs(int64_t n, const int64_t* colptr, const int32_t* rows,
                      const uint8_t* skip, int64_t* pairs_out) {
  if (n > 128) return -1;
  /* r[i][j+1] = rank(rows >= i, cols <= j); r[n][*] = 0; r[*][0] = 0 */
  int64_t S = n + 1;
  int32_t* r = (int32_t*)calloc((size_t)(S * S), sizeof(int32_t));
  for (int64_t i = 0; i < n; i++) {
    uint64_t basis[128][2];
    int32_t bl[128];
    int nb = 0;
    int32_t rank = 0;
    for (int64_t j = 0; j < n; j++) {
      uint64_t v[2] = {0, 0};
      if (!(skip && skip[j]))
        for (int64_t k = colptr[j]; k < colptr[j + 1]; k++)
          if (rows[k] >= i) v[rows[k] >> 6] |= 1ULL << (rows[k] & 63);
      for (;;) {
        int32_t hb = v[1] ? 64 + 63 - __builtin_clzll(v[1]) : (v[0] ? 63 - __builtin_clzll(v[0]) : -1);
        if (hb < 0) break;
        int found = -1;
        for (int b = 0; b < nb; b++)
          if (bl[b] == hb) { found = b; break; }
        if (found < 0) {
          basis[nb][0] = v[0];
          basis[nb][1] = v[1];
          bl[nb++] = hb;
          rank++;
          break;
        }
        v[0] ^= basis[found][0];
        v[1] ^= basis[found][1];
      }
      r[i * S + j + 1] = rank;
    }
  }
  int32_t* dob = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; i++) dob[i] = -1;
  for (int64_t i = 0; i < n; i++)
    for (int64_t j = 0; j < n; j++) {
      int32_t a = r[i * S + j + 1], b = r[(i + 1) * S + j + 1], c = r[i * S + j], d = r[(i + 1) * S + j];
      if (a - b - c + d == 1) dob[i] = (int32_t)j;
    }
  int64_t np = emit_pairs(n, dob, pairs_out);
  free(r);
  free(dob);
  return np;
}

/* ------------------------------------------------------------------------------------ */
/* GUDHI edge collapse (Flag_complex_edge_collapser.h:98-524), restated.                  */
/*                                                                                        */
/* Edge e has filtration e.  An edge is "valid" for a test at index curr when e < curr or  */
/* e is critical (valid_it_open, :148-152); a vertex is in its own closed neighbourhood    */
/* (diagonal entry, :227-231).  edge_is_dominated (:224-303) is the predicate             */
/*   I = common valid neighbours of (a,b);  |I|==1  or  exists w in I with I in N[w].       */
/* process_edges (:416-482): edge 0 is critical; edge i is tested against the graph of     */
/* edges < i; a critical edge triggers set_edge_critical (:326-359), which visits the      */
/* non-critical edges of its 3-cliques in decreasing index order, re-tests each at its own */
/* index and, when it is no longer dominated, makes it critical and enqueues its 3-cliques.*/
/* ------------------------------------------------------------------------------------ */
typedef struct {
  int64_t nv;
  int64_t* off;   /* nv+1 */
  int32_t* nbr;   /* sorted neighbours */
  int32_t* eid;   /* edge index of each adjacency entry */
  uint8_t* crit;  /* per edge */
  int32_t* eu;
  int32_t* ev;
} cgraph;

static inline int valid_open(const cgraph* g, int32_t e, int32_t curr) {
  return e < curr || g->crit[e];
}
/* edge index of (a,x) or -1 */
static int32_t find_e(const cgraph* g, int32_t a, int32_t x) {
  int64_t lo = g->off[a], hi = g->off[a + 1];
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (g->nbr[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return (lo < g->off[a + 1] && g->nbr[lo] == x) ? g->eid[lo] : -1;
}
/* common valid neighbours of a and b at curr -> out (sorted); returns count */
static int64_t common_valid(const cgraph* g, int32_t a, int32_t b, int32_t curr, int32_t* out,
                            int32_t* ea, int32_t* eb) {
  int64_t i = g->off[a], ie = g->off[a + 1], j = g->off[b], je = g->off[b + 1], k = 0;
  while (i < ie && j < je) {
    if (!valid_open(g, g->eid[i], curr)) { i++; continue; }
    if (!valid_open(g, g->eid[j], curr)) { j++; continue; }
    if (g->nbr[i] < g->nbr[j]) i++;
    else if (g->nbr[i] > g->nbr[j]) j++;
    else {
      if (ea) ea[k] = g->eid[i];
      if (eb) eb[k] = g->eid[j];
      out[k++] = g->nbr[i];
      i++;
      j++;
    }
  }
  return k;
}
static int dominated(const cgraph* g, int32_t a, int32_t b, int32_t curr, int32_t* I) {
  int64_t m = common_valid(g, a, b, curr, I, NULL, NULL);
  if (m == 0) return 0;
  if (m == 1) return 1;
  for (int64_t c = 0; c < m; c++) {
    int32_t w = I[c];
    /* I subset of closed valid neighbourhood of w: merge-walk I against N(w) */
    int64_t p = g->off[w], pe = g->off[w + 1];
    int ok = 1;
    for (int64_t t = 0; t < m && ok; t++) {
      int32_t x = I[t];
      if (x == w) continue;
      while (p < pe && g->nbr[p] < x) p++;
      if (p >= pe || g->nbr[p] != x || !valid_open(g, g->eid[p], curr)) ok = 0;
    }
    if (ok) return 1;
  }
  return 0;
}

typedef struct {
  int32_t* h;
  int64_t n, cap;
} maxheap;
static void hp_push(maxheap* hp, int32_t v) {
  if (hp->n == hp->cap) {
    hp->cap = hp->cap ? hp->cap * 2 : 256;
    hp->h = (int32_t*)realloc(hp->h, sizeof(int32_t) * (size_t)hp->cap);
  }
  int64_t i = hp->n++;
  hp->h[i] = v;
  while (i > 0) {
    int64_t p = (i - 1) / 2;
    if (hp->h[p] >= hp->h[i]) break;
    int32_t t = hp->h[p];
    hp->h[p] = hp->h[i];
    hp->h[i] = t;
    i = p;
  }
}
static int32_t hp_pop(maxheap* hp) {
  int32_t top = hp->h[0];
  hp->h[0] = hp->h[--hp->n];
  int64_t i = 0;
  for (;;) {
    int64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < hp->n && hp->h[l] > hp->h[m]) m = l;
    if (r < hp->n && hp->h[r] > hp->h[m]) m = r;
    if (m == i) break;
    int32_t t = hp->h[m];
    hp->h[m] = hp->h[i];
    hp->h[i] = t;
    i = m;
  }
  return top;
}

static void three_clique(const cgraph* g, int32_t thresh, int32_t a, int32_t b, int32_t* I,
                         int32_t* ea, int32_t* eb, maxheap* hp, int32_t* stamp, int32_t sid) {
  int64_t m = common_valid(g, a, b, thresh, I, ea, eb);
  for (int64_t k = 0; k < m; k++) {
    int32_t e1 = ea[k], e2 = eb[k];
    if (!g->crit[e1] && e1 < thresh && stamp[e1] != sid) { stamp[e1] = sid; hp_push(hp, e1); }
    if (!g->crit[e2] && e2 < thresh && stamp[e2] != sid) { stamp[e2] = sid; hp_push(hp, e2); }
  }
}

int or_collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept) {
  if (ne <= 0) return 0;
  /* compact vertex labels */
  int32_t* lab = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * ne));
  for (int64_t i = 0; i < ne; i++) { lab[2 * i] = u[i]; lab[2 * i + 1] = v[i]; }
  int64_t nv = sort_unique_i32(lab, 2 * ne);
  cgraph g;
  memset(&g, 0, sizeof(g));
  g.nv = nv;
  g.eu = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  g.ev = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  for (int64_t i = 0; i < ne; i++) {
    int32_t* pu = (int32_t*)bsearch(&u[i], lab, (size_t)nv, 4, cmp_i32);
    int32_t* pv = (int32_t*)bsearch(&v[i], lab, (size_t)nv, 4, cmp_i32);
    g.eu[i] = (int32_t)(pu - lab);
    g.ev[i] = (int32_t)(pv - lab);
  }
  g.off = (int64_t*)calloc((size_t)nv + 1, sizeof(int64_t));
  for (int64_t i = 0; i < ne; i++) { g.off[g.eu[i] + 1]++; g.off[g.ev[i] + 1]++; }
  for (int64_t i = 0; i < nv; i++) g.off[i + 1] += g.off[i];
  g.nbr = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * ne));
  g.eid = (int32_t*)malloc(sizeof(int32_t) * (size_t)(2 * ne));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)nv);
  memcpy(fill, g.off, sizeof(int64_t) * (size_t)nv);
  for (int64_t i = 0; i < ne; i++) {
    int64_t p = fill[g.eu[i]]++;
    g.nbr[p] = g.ev[i]; g.eid[p] = (int32_t)i;
    p = fill[g.ev[i]]++;
    g.nbr[p] = g.eu[i]; g.eid[p] = (int32_t)i;
  }
  /* sort each adjacency row by neighbour (insertion sort on pairs; rows are short) */
  for (int64_t a = 0; a < nv; a++) {
    for (int64_t i = g.off[a] + 1; i < g.off[a + 1]; i++) {
      int32_t xn = g.nbr[i], xe = g.eid[i];
      int64_t j = i - 1;
      while (j >= g.off[a] && g.nbr[j] > xn) { g.nbr[j + 1] = g.nbr[j]; g.eid[j + 1] = g.eid[j]; j--; }
      g.nbr[j + 1] = xn; g.eid[j + 1] = xe;
    }
  }
  g.crit = (uint8_t*)calloc((size_t)ne, 1);
  int64_t maxdeg = 1;
  for (int64_t a = 0; a < nv; a++)
    if (g.off[a + 1] - g.off[a] > maxdeg) maxdeg = g.off[a + 1] - g.off[a];
  int32_t* I = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxdeg);
  int32_t* ea = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxdeg);
  int32_t* eb = (int32_t*)malloc(sizeof(int32_t) * (size_t)maxdeg);
  int32_t* stamp = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  for (int64_t i = 0; i < ne; i++) stamp[i] = -1;
  maxheap hp = {NULL, 0, 0};

  memset(kept, 0, (size_t)ne);
  g.crit[0] = 1; /* the first edge is always critical (:449-462) */
  kept[0] = 1;
  for (int32_t idx = 1; idx < (int32_t)ne; idx++) {
    if (dominated(&g, g.eu[idx], g.ev[idx], idx, I)) continue;
    g.crit[idx] = 1;
    kept[idx] = 1;
    /* set_edge_critical(idx) */
    hp.n = 0;
    three_clique(&g, idx, g.eu[idx], g.ev[idx], I, ea, eb, &hp, stamp, idx);
    while (hp.n) {
      int32_t cur = hp_pop(&hp);
      if (g.crit[cur]) continue;
      if (dominated(&g, g.eu[cur], g.ev[cur], cur, I)) continue;
      g.crit[cur] = 1;
      kept[cur] = 1;
      three_clique(&g, cur, g.eu[cur], g.ev[cur], I, ea, eb, &hp, stamp, idx);
    }
  }
  free(hp.h);
  free(stamp);
  free(I);
  free(ea);
  free(eb);
  free(fill);
  free(g.off);
  free(g.nbr);
  free(g.eid);
  free(g.crit);
  free(g.eu);
  free(g.ev);
  free(lab);
  return 0;
}

int64_t or_apply_collapse(int64_t n, const int64_t* colptr, const int32_t* rows,
                          const int32_t* dims, uint8_t* skip) {
  int64_t ne = 0;
  int32_t maxd = 0;
  if (n < 0) return -1;
  for (int64_t c = 0; c < n; c++) {
    if (dims[c] == 1) {
      if (colptr[c + 1] - colptr[c] != 2) return -1;
      ne++;
    }
    if (dims[c] > maxd) maxd = dims[c];
  }
  if (n > 0) memset(skip, 0, (size_t)n);
  if (ne == 0) return 0;
  int32_t* u = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  int32_t* v = (int32_t*)malloc(sizeof(int32_t) * (size_t)ne);
  int64_t* col = (int64_t*)malloc(sizeof(int64_t) * (size_t)ne);
  uint8_t* kept = (uint8_t*)malloc((size_t)ne);
  int64_t e = 0;
  for (int64_t c = 0; c < n; c++)
    if (dims[c] == 1) {
      u[e] = rows[colptr[c]];
      v[e] = rows[colptr[c] + 1];
      col[e++] = c;
    }
  or_collapse_edges(ne, u, v, kept);
  int64_t nk = 0;
  for (int64_t i = 0; i < ne; i++) {
    if (!kept[i]) skip[col[i]] = 1;
    else nk++;
  }
  /* coface propagation, one dimension at a time (kernel.h:190-198) */
  for (int32_t d = 2; d <= maxd; d++)
    for (int64_t c = 0; c < n; c++) {
      if (dims[c] != d) continue;
      for (int64_t k = colptr[c]; k < colptr[c + 1]; k++)
        if (skip[rows[k]]) { skip[c] = 1; break; }
    }
  free(u);
  free(v);
  free(col);
  free(kept);
  return nk;
}
