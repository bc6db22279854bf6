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

}  // namespace ssq
