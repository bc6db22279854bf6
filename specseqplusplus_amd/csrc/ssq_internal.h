// Internal declarations shared by the libssq translation units.
#pragma once
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

}  // namespace ssq
