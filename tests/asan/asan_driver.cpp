// Host-side sanitizer driver (AddressSanitizer + UndefinedBehaviorSanitizer) for the CPU parts
// of libssq (PHAT reader/writer, host Rips builder, edge collapse) and for the oracle.  It
// exercises well-formed and malformed inputs; any sanitizer report fails the run.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ssq.h"
#include "../../oracle/ssq_oracle.h"

static int fails = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      fails++;                                                      \
    }                                                               \
  } while (0)

static std::string write_file(const std::string& dir, const char* name, const void* data, size_t n) {
  const std::string p = dir + "/" + name;
  FILE* f = std::fopen(p.c_str(), "wb");
  std::fwrite(data, 1, n, f);
  std::fclose(f);
  return p;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  // ---- PHAT reader: good and malformed ASCII / binary ----
  const char good[] = "# comment\n0\n0\n0\n1 0 1\n1 0 2\n1 1 2\n\n2 3 4 5\n";
  // rejected: row past the end, bad token, negative row, absurd dimension; accepted (as the
  // reference reader does): repeated rows (sorted and de-duplicated), a row that is in range
  const char* bad_ascii[] = {"0\n1 0 7\n", "0\n1 x 0\n", "1 -1 0\n", "9999999999999 1\n"};
  const char* odd_ascii[] = {"0\n0\n1 1 1 0\n", "3 0\n"};
  const int64_t good_bin[] = {7, 0, 0, 0, 0, 0, 0, 1, 2, 0, 1, 1, 2, 0, 2, 1, 2, 1, 2, 2, 3, 3, 4, 5};
  const int64_t bad_bin1[] = {3, 0, 0, 1, 2, 0, 5};             // row out of range
  const int64_t bad_bin2[] = {4, 0, 0, 1, 1000000000000LL, 0};  // absurd k
  const int64_t bad_bin3[] = {2, 0};                            // truncated
  ssq_matrix m{};
  EXPECT(ssq_load_phat(write_file(dir, "g.dat", good, sizeof(good) - 1).c_str(), 0, &m) == SSQ_OK);
  EXPECT(m.n == 7 && m.nnz == 9);
  ssq_matrix_free(&m);
  EXPECT(ssq_load_phat(write_file(dir, "g.bin", good_bin, sizeof(good_bin)).c_str(), 1, &m) == SSQ_OK);
  EXPECT(m.n == 7);
  ssq_matrix_free(&m);
  for (size_t i = 0; i < sizeof(bad_ascii) / sizeof(bad_ascii[0]); i++) {
    std::memset(&m, 0, sizeof(m));
    EXPECT(ssq_load_phat(write_file(dir, "b.dat", bad_ascii[i], std::strlen(bad_ascii[i])).c_str(), 0, &m) != SSQ_OK);
    ssq_matrix_free(&m);
  }
  const std::pair<const int64_t*, size_t> bad_bins[] = {
      {bad_bin1, sizeof(bad_bin1)}, {bad_bin2, sizeof(bad_bin2)}, {bad_bin3, sizeof(bad_bin3)}};
  for (const auto& b : bad_bins) {
    std::memset(&m, 0, sizeof(m));
    EXPECT(ssq_load_phat(write_file(dir, "b.bin", b.first, b.second).c_str(), 1, &m) != SSQ_OK);
    ssq_matrix_free(&m);
  }
  for (size_t i = 0; i < sizeof(odd_ascii) / sizeof(odd_ascii[0]); i++) {
    std::memset(&m, 0, sizeof(m));
    EXPECT(ssq_load_phat(write_file(dir, "o.dat", odd_ascii[i], std::strlen(odd_ascii[i])).c_str(), 0, &m) == SSQ_OK);
    ssq_matrix_free(&m);
  }
  EXPECT(ssq_load_phat((dir + "/missing").c_str(), 0, &m) == SSQ_EIO);
  // oracle reader on the same inputs
  int64_t on = 0;
  int64_t* ocp = nullptr;
  int32_t *orw = nullptr, *odm = nullptr;
  EXPECT(or_load_phat((dir + "/g.dat").c_str(), 0, &on, &ocp, &orw, &odm) == 0 && on == 7);
  std::vector<int64_t> pr(2 * 7);
  EXPECT(or_twist_reduce(on, ocp, orw, odm, nullptr, pr.data(), nullptr) == 3);
  EXPECT(or_twist_reduce_dual(on, ocp, orw, odm, nullptr, pr.data(), nullptr) == 3);
  EXPECT(or_dense_standard(on, ocp, orw, nullptr, pr.data()) == 3);
  or_free(ocp);
  or_free(orw);
  or_free(odm);
  for (size_t i = 0; i < sizeof(bad_ascii) / sizeof(bad_ascii[0]); i++) {
    write_file(dir, "b.dat", bad_ascii[i], std::strlen(bad_ascii[i]));
    ocp = nullptr;
    orw = nullptr;
    odm = nullptr;
    EXPECT(or_load_phat((dir + "/b.dat").c_str(), 0, &on, &ocp, &orw, &odm) != 0);
  }
  // ---- pair writer ----
  int64_t pp[6] = {1, 3, 2, 4, 5, 6};
  ssq_pairs P{3, pp, 0, nullptr};
  EXPECT(ssq_save_pairs((dir + "/p.dat").c_str(), 0, &P) == SSQ_OK);
  EXPECT(ssq_save_pairs((dir + "/p.bin").c_str(), 1, &P) == SSQ_OK);
  EXPECT(or_save_pairs((dir + "/q.dat").c_str(), 0, 3, pp) == 0);
  // ---- host Rips builder, edge collapse, oracle reductions on random clouds ----
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<float> U(0.f, 1.f);
  for (int t = 0; t < 6; t++) {
    const int n = 10 + 15 * t, pdim = 2 + t % 3;
    std::vector<float> X((size_t)n * pdim);
    for (auto& x : X) x = U(rng);
    if (t == 5) X.assign(X.size(), 0.5f);  // every point repeated
    ssq_matrix v{};
    EXPECT(ssq_vr_build(X.data(), n, pdim, 0.25f + 0.1f * t, 2 + (t & 1), &v) == SSQ_OK);
    std::vector<int64_t> out(2 * (size_t)v.n + 2);
    const int64_t a = or_twist_reduce(v.n, v.col_ptr, v.rows, v.dims, nullptr, out.data(), nullptr);
    const int64_t b = or_twist_reduce_dual(v.n, v.col_ptr, v.rows, v.dims, nullptr, out.data(), nullptr);
    EXPECT(a == b);
    std::vector<uint8_t> skip((size_t)v.n, 0);
    EXPECT(or_apply_collapse(v.n, v.col_ptr, v.rows, v.dims, skip.data()) >= 0);
    // edges in filtration order for the collapsers
    std::vector<int32_t> eu, ev;
    for (int64_t c = 0; c < v.n; c++)
      if (v.dims[c] == 1) {
        eu.push_back(v.rows[v.col_ptr[c]]);
        ev.push_back(v.rows[v.col_ptr[c] + 1]);
      }
    std::vector<uint8_t> k1(eu.size() + 1), k2(eu.size() + 1);
    EXPECT(ssq_collapse_edges((int64_t)eu.size(), eu.data(), ev.data(), k1.data(), 4) == SSQ_OK);
    EXPECT(or_collapse_edges((int64_t)eu.size(), eu.data(), ev.data(), k2.data()) == 0);
    EXPECT(std::memcmp(k1.data(), k2.data(), eu.size()) == 0);
    ssq_matrix_free(&v);
  }
  ssq_matrix bad{};
  EXPECT(ssq_vr_build(nullptr, 5, 3, 0.1f, 2, &bad) != SSQ_OK);
  // alpha complexes: random points, a cospherical grid with duplicates, and refused inputs
  for (int t = 0; t < 3; t++) {
    std::vector<double> P;
    if (t < 2) {
      for (int i = 0; i < 3 * (60 + 90 * t); i++) P.push_back(t ? U(rng) - 0.5 : U(rng));  // signs both ways
    } else {
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++)
          for (int z = 0; z < 4; z++) P.insert(P.end(), {0.5 * x, 0.5 * y, 0.5 * z});
      P.insert(P.end(), P.begin(), P.begin() + 12);  // four duplicates
    }
    const int64_t n = (int64_t)P.size() / 3;
    ssq_matrix a{};
    double* vals = nullptr;
    EXPECT(ssq_alpha_build(P.data(), n, t == 1 ? -1.0 : 0.05, t & 1, &a, &vals) == SSQ_OK);
    for (int64_t c = 0; c < a.n; c++)
      for (int64_t k = a.col_ptr[c]; k < a.col_ptr[c + 1]; k++) EXPECT(a.rows[k] < c);
    for (int64_t c = 1; c < a.n && !(t & 1); c++) EXPECT(vals[c] >= vals[c - 1]);
    std::vector<int32_t> tets(4 * 8 * (size_t)n + 64);
    int64_t nt = 0;
    EXPECT(ssq_delaunay_3(P.data(), n, tets.data(), (int64_t)tets.size() / 4, &nt) == SSQ_OK && nt > 0);
    ssq_free(vals);
    ssq_matrix_free(&a);
  }
  {
    std::vector<double> flat;  // coplanar points: refused
    for (int i = 0; i < 20; i++) flat.insert(flat.end(), {U(rng), U(rng), 0.0});
    ssq_matrix a{};
    EXPECT(ssq_alpha_build(flat.data(), 20, 1.0, 0, &a, nullptr) == SSQ_EINVAL);
    std::vector<double> nan_pts(12, 0.25);
    nan_pts[5] = std::nan("");
    EXPECT(ssq_alpha_build(nan_pts.data(), 4, 1.0, 0, &a, nullptr) == SSQ_EINVAL);
    EXPECT(ssq_alpha_build(nan_pts.data(), 3, 1.0, 0, &a, nullptr) == SSQ_EINVAL);
  }
  std::printf("asan driver: %d failed checks\n", fails);
  return fails ? 1 : 0;
}
