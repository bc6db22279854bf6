// specseq++ command line: the same 12 positional arguments, help text and exit codes as the
// reference (main.cu:269-352; README.md:26), on top of the libssq C ABI.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ssq.h"

static void print_help() {
  std::printf(
      "Usage: specseq++ <input_file> <output_file> <input_format> <output_format> <dualize> <model> "
      "<max_iteration> <ss_block_size> <block_size_growth_rate> <queue_size> <level1_capacity> "
      "<level2_capacity>\n"
      "input_format: ascii or binary\n"
      "output_format: ascii or binary\n"
      "dualize: \n  0 - No dualize\n  1 - Enable dualize\n"
      "model: \n  0 - No optimizations\n  1 - Enable high-dimensional clearing theorem\n"
      "  2 - Enable edge collapsing\n"
      "  3 - Enable both high-dimension guided clearing theorem and edge collapsing\n"
      "max_iteration[Only effective when model equals 1 or 3.]: neighbours examined per max-dimension cell by "
      "the guided clearing.\n"
      "ss_block_size: Block size for ss (default 102400)\n"
      "block_size_growth_rate: Block size growth rate (default 1.05)\n"
      "queue_size: maximum size addition (default 65536)\n"
      "level1_capacity: Capacity for level 1 array\n"
      "level2_capacity: Capacity for level 2 array\n"
      "Example: ./specseq++ input.txt output.txt ascii ascii 0 1 10 1024000 1.05 65536 1000 2000\n");
}

static bool parse_int(const char* s, int32_t* out) {
  char* e = nullptr;
  long v = std::strtol(s, &e, 10);
  if (e == s) return false;
  *out = (int32_t)v;
  return true;
}

int main(int argc, char** argv) {
  if (argc == 2 && (!std::strcmp(argv[1], "-help") || !std::strcmp(argv[1], "-h"))) {
    print_help();
    return 0;
  }
  if (argc < 13) {
    std::fprintf(stderr, "Invalid number of arguments.\n");
    print_help();
    return 1;
  }
  ssq_params p;
  ssq_default_params(&p);
  int in_bin, out_bin;
  if (!std::strcmp(argv[3], "ascii")) in_bin = 0;
  else if (!std::strcmp(argv[3], "binary")) in_bin = 1;
  else {
    std::fprintf(stderr, "Invalid input format. Format should be 'ascii' or 'binary'.\n");
    print_help();
    return 1;
  }
  if (!std::strcmp(argv[4], "ascii")) out_bin = 0;
  else if (!std::strcmp(argv[4], "binary")) out_bin = 1;
  else {
    std::fprintf(stderr, "Invalid output format. Format should be 'ascii' or 'binary'.\n");
    print_help();
    return 1;
  }
  char* e = nullptr;
  if (!parse_int(argv[5], &p.dualize) || !parse_int(argv[6], &p.model) || !parse_int(argv[7], &p.max_iteration) ||
      !parse_int(argv[8], &p.ss_block_size) || !parse_int(argv[10], &p.queue_size) ||
      !parse_int(argv[11], &p.level1) || !parse_int(argv[12], &p.level2)) {
    std::fprintf(stderr, "Invalid numeric argument.\n");
    print_help();
    return 1;
  }
  p.growth = std::strtof(argv[9], &e);
  if (e == argv[9]) {
    std::fprintf(stderr, "Invalid numeric argument.\n");
    print_help();
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  ssq_matrix m;
  int rc = ssq_load_phat(argv[1], in_bin, &m);
  if (rc) {
    std::fprintf(stderr, "Error: Failed to read %s (status %d).\n", argv[1], rc);
    return 2;
  }
  auto t1 = std::chrono::steady_clock::now();
  std::printf("-----------------------------------------------Reading  completed-----------------------------------------------\n");
  ssq_handle* h = nullptr;
  rc = ssq_create(0, &h);
  if (rc) {
    std::fprintf(stderr, "Error: no usable HIP device (status %d).\n", rc);
    ssq_matrix_free(&m);
    return 3;
  }
  ssq_pairs pr;
  ssq_stats st;
  rc = ssq_reduce(h, &m, &p, &pr, &st);
  if (rc) {
    std::fprintf(stderr, "Error: reduction failed (status %d): %s\n", rc, ssq_last_error(h));
    ssq_destroy(h);
    ssq_matrix_free(&m);
    return 4;
  }
  std::printf("NumCols: %lld  pending after pre-passes: %lld  apparent pairs: %lld  guided-cleared: %lld  kept edges: %lld\n",
              (long long)m.n, (long long)st.n_pending, (long long)st.n_apparent, (long long)st.n_cleared_hd,
              (long long)st.n_kept_edges);
  std::printf("pages: %lld  rounds: %lld  additions: %lld\n", (long long)st.n_pages, (long long)st.n_rounds,
              (long long)st.n_additions);
  std::printf("Initialization Time: %10.6f seconds\n", st.t_init_ms / 1000.0);
  std::printf("Computation    Time: %10.6f seconds\n", st.t_reduce_ms / 1000.0);
  std::printf("Total          Time: %10.6f seconds\n", st.t_total_ms / 1000.0);
  std::printf("Read           Time: %10.6f seconds\n", std::chrono::duration<double>(t1 - t0).count());
  rc = ssq_save_pairs(argv[2], out_bin, &pr);
  if (rc) {
    std::fprintf(stderr, "Error: Could not write %s (status %d).\n", argv[2], rc);
    ssq_pairs_free(&pr);
    ssq_destroy(h);
    ssq_matrix_free(&m);
    return 5;
  }
  std::printf("-----------------------------------------------Writing  completed-----------------------------------------------\n");
  ssq_pairs_free(&pr);
  ssq_destroy(h);
  ssq_matrix_free(&m);
  return 0;
}
