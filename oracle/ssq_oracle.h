/*
 * ssq oracle — CPU restatement of the SpecSeq++ persistent-homology reduction path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker (or as the timed CPU
 * baseline). The product library (libssq) never links, loads or calls it.
 *
 * Each function names the reference code it restates and how it is pinned:
 *
 *  or_load_phat      main.cu:90-167 (PHAT ASCII/binary reader; '#'/blank lines skipped,
 *                    rows sorted + de-duplicated).  Pinned by the single_triangle KAT
 *                    (phat-highclear/examples/single_triangle.{dat,bin}) in ASCII and binary.
 *  or_save_pairs     main.cu:191-229 (count line, then "birth death" lines; binary int64).
 *                    Pinned byte-for-byte by single_triangle_persistence_pairs.{dat,bin}.
 *  or_twist_reduce   phat.cpp:97-164,184-186 (PHAT twist_reduction, the reference's CPU
 *                    oracle) and specseq.h:592-1060,1182-1193 (twist order, pair mapping).
 *                    Pinned by: the single_triangle KAT; agreement with or_dense_standard
 *                    (standard algorithm, no clearing) and with or_rank_pairs (pivots from
 *                    ranks of lower-left submatrices, an algorithm-free characterisation) on
 *                    random tiny complexes and with or_dense_standard on the reference's
 *                    shipped 20,875-cell data/ matrices; Betti numbers of triangulated spheres /
 *                    tori / projective plane (closed forms).  (The shipped Alpha *_maxclear files
 *                    are byte-identical to their originals and add no pin of their own.)
 *  or_twist_reduce_dual  specseq.h:303-363,1188 (dualize: twist reduction of the
 *                    anti-transpose, pairs mapped back as (b', d') -> (n-1-d', n-1-b')).
 *                    Pinned by agreement with or_twist_reduce (persistence duality: homology
 *                    and cohomology pairs coincide) and with or_rank_pairs on random tiny
 *                    complexes, closed forms and the shipped fixtures.
 *  or_dense_standard standard reduction (Edelsbrunner-Harer p.153) on dense bit columns.
 *                    Pinned by or_rank_pairs on tiny inputs.
 *  or_rank_pairs     (i,j) is a pair iff r(i,j)-r(i+1,j)-r(i,j-1)+r(i+1,j-1)=1, with r the
 *                    GF(2) rank of rows>=i, cols<=j.  Textbook closed form; no reduction.
 *  or_collapse_edges Flag_complex_edge_collapser.h:98-524 (GUDHI sequential edge collapse),
 *                    as called by specseq.h:209-212,410-424 (edge filtration = edge rank).
 *                    Pinned against oracle/_ref (the reference header compiled from its own
 *                    source) and the committed kept-edge fixtures under tests/golden/.
 *  or_apply_collapse specseq.h:417-423 + kernel.h:190-198 (non-kept edges and all their
 *                    cofaces are emptied).  Pinned by the sub-complex invariants and by
 *                    homotopy invariance: a strong (edge) collapse keeps the flag complex's
 *                    homotopy type, so the kept sub-complex has the full complex's beta_0 and
 *                    beta_1 (the 2-skeleton's beta_2 need not agree); a planted over-removal
 *                    fails it.  (The kept edges' changed critical times that GUDHI reports are
 *                    dropped by the reference, specseq.h:421-423, so bars are not preserved.)
 */
#ifndef SSQ_ORACLE_H
#define SSQ_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Matrix arrays are malloc'd by the loader; free each with or_free. Returns 0 on success,
 * 1 = cannot open, 2 = format error (index out of range / bad token / truncated). */
int or_load_phat(const char* path, int binary, int64_t* n_out, int64_t** colptr_out,
                 int32_t** rows_out, int32_t** dims_out);
void or_free(void* p);

/* pairs: np (birth, death) int64 pairs. Returns 0 on success. */
int or_save_pairs(const char* path, int binary, int64_t np, const int64_t* pairs);

/* Twist reduction with clearing. skip[c]!=0 empties column c beforehand (sub-complex).
 * pairs_out must hold 2*n int64. Returns the number of pairs (sorted by birth).
 * stats_out (may be NULL, else 4 doubles) receives {additions, algorithmic bytes, most
 * additions made by one column, dependency depth} where one addition a <- a xor b costs
 * 4*(|a|+|b|+|a xor b|)+24 bytes, and the dependency depth is the sum over dimensions of the
 * longest chain of additions when each addition takes one step and may start once the pivot
 * it adds is final (the time of an unlimited-parallel reduction, in additions). */
int64_t or_twist_reduce(int64_t n, const int64_t* colptr, const int32_t* rows,
                        const int32_t* dims, const uint8_t* skip, int64_t* pairs_out,
                        double* stats_out);

/* The same pairs by twist reduction of the anti-transpose (cohomology), mapped back.  Same
 * output contract; stats_out counts the dual reduction's additions. */
int64_t or_twist_reduce_dual(int64_t n, const int64_t* colptr, const int32_t* rows,
                             const int32_t* dims, const uint8_t* skip, int64_t* pairs_out,
                             double* stats_out);

/* Standard algorithm on dense bit columns (n <= 65536). Same output contract. */
int64_t or_dense_standard(int64_t n, const int64_t* colptr, const int32_t* rows,
                          const uint8_t* skip, int64_t* pairs_out);

/* Rank-formula pairs (n <= 96). Same output contract. */
int64_t or_rank_pairs(int64_t n, const int64_t* colptr, const int32_t* rows,
                      const uint8_t* skip, int64_t* pairs_out);

/* GUDHI edge collapse over edges given in filtration order (edge i has filtration i).
 * u[i], v[i] are vertex ids (any labels).  kept[i] = 1 for every edge GUDHI outputs. */
int or_collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept);

/* Model-2 sub-complex: marks skip[c]=1 for every dim-1 column that the collapse does not
 * keep and every coface of a marked column.  Returns the number of kept edges. */
int64_t or_apply_collapse(int64_t n, const int64_t* colptr, const int32_t* rows,
                          const int32_t* dims, uint8_t* skip);

#ifdef __cplusplus
}
#endif
#endif
