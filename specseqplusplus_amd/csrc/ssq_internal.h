// Internal declarations shared by the libssq translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ssq.h"

namespace ssq {

int matrix_from_vectors(const std::vector<int64_t>& colptr, const std::vector<int32_t>& rows,
                        const std::vector<int32_t>& dims, ssq_matrix* out);

// Flag-complex edge collapse with the exact kept-edge set of the sequential algorithm in
// Flag_complex_edge_collapser.h:416-482 (edge i has filtration i).  kept[i] = 1 for kept.
void collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept, int threads);

// Vietoris-Rips boundary matrix (host builder).
int vr_build(const float* pts, int64_t npts, int pdim, float r, int maxdim, ssq_matrix* out);
// Its edge and triangle counts (adjacency bitsets; no cells are built).
void rips_count_host(const float* pts, int64_t npts, int pdim, float r, int64_t* edges, int64_t* triangles);

// GPU Vietoris-Rips builder (dimension <= 2).  Device outputs stay valid until the next
// build on the same builder.
struct VrBuilder;
VrBuilder* vr_builder_create();
void vr_builder_destroy(VrBuilder* b);
int vr_builder_build(VrBuilder* b, hipStream_t s, const float* h_pts, int64_t npts, int pdim, float r, int maxdim,
                     const int64_t** d_colptr, const int32_t** d_rows, const int32_t** d_dims, int64_t* n_out,
                     int64_t* nnz_out);
// Many clouds (same pdim) in one build: the output is their block-diagonal matrix, cloud c's
// cells in rows/columns [cell_base[c], cell_base[c + 1]) in its own (filtration, dim, lex)
// order.  cell_base (count + 1 entries) may be null.
int vr_builder_build_batch(VrBuilder* b, hipStream_t s, int count, const float* const* h_pts, const int64_t* npts,
                           int pdim, const float* radius, int maxdim, const int64_t** d_colptr, const int32_t** d_rows,
                           const int32_t** d_dims, int64_t* n_out, int64_t* nnz_out, int64_t* cell_base);

// What the reduction engine offers a builder that writes a matrix straight into its buffers
// (Engine::prebuilt_begin / prebuilt_reduce in ssq_engine.hip).
struct PrebuiltTarget {
  virtual ~PrebuiltTarget() = default;
  // cells [0, n) are rows, [0, ncols) also columns (colptr[ncols + 1], dims[ncols]); returns
  // pool (nnz rows, then arena free entries), leftone[n] (to fill) and negflag[n] (zeroed)
  virtual int pre_begin(int64_t n, int64_t ncols, int64_t nnz, int64_t arena, const int64_t* colptr,
                        const int32_t* dims, int32_t maxdim, int32_t** pool, uint32_t** leftone,
                        uint8_t** negflag) = 0;
  virtual int pre_reduce(const ssq_params& p, ssq_stats* st, bool* oom) = 0;
  virtual const int32_t* pre_low() const = 0;
  virtual const uint32_t* pre_piv() const = 0;
  virtual hipStream_t pre_stream() const = 0;
  virtual int pre_device() const = 0;
  virtual const std::string& pre_error() const = 0;
};

// Batched Vietoris-Rips cohomology (dimension <= 2, clouds of at most kRipsMaxPoints points)
// built straight into the engine's dual layout (ssq_ripsco.hip).
constexpr int kRipsMaxPoints = 4096;
struct RipsCo;
RipsCo* ripsco_create();
void ripsco_destroy(RipsCo* b);
// Exact edge and triangle counts of each cloud (device points, host bases and thresholds).
int ripsco_count(RipsCo* b, hipStream_t s, int count, const float* d_pts, const int64_t* pbase, int pdim,
                 const float* radius, int64_t* edges, int64_t* triangles);
// Pairs of every cloud, sorted by birth, in the indices of its own (filtration, dim, lex)
// ordered complex: d_pairs[2 * pair_off[c] ...] for cloud c (pair_off: count + 1 entries).
// Clouds are reduced in groups sized to the free device memory (mem_budget 0 = automatic,
// shared by the streams).
// With d_pairs null the pairs go to a buffer of the builder's, returned in *d_pairs_owned
// (valid until its next call).
// Groups run on nstreams engines side by side (builder b[k] with engine eng[k]; b[0] keeps the
// sizing pass and the output).
int ripsco_reduce(RipsCo* const* b, PrebuiltTarget* const* eng, int nstreams, int count, const float* d_pts,
                  const int64_t* pbase, int pdim, const float* radius, const ssq_params& p, int64_t* d_pairs,
                  int64_t pairs_cap, int64_t* pair_off, int64_t mem_budget, ssq_stats* st, std::string* err,
                  const int64_t** d_pairs_owned);
// Streams the batched Rips path of this handle uses (a server worker uses one: its workers
// already run side by side).
void handle_set_rips_streams(ssq_handle* h, int n);

// Triangle count of each cloud from the last ripsco_count / ripsco_reduce call.
const int64_t* ripsco_triangles(const RipsCo* b);
const int64_t* ripsco_edges(const RipsCo* b);

}  // namespace ssq
