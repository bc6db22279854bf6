/*
 * libssq — MI355X (gfx950) persistent-homology reduction engine, C ABI.
 *
 * Drop-in boundary for the SpecSeq++ reduction path. Every entry point takes plain
 * pointers and sizes, returns an int status (never exits), and names the reference
 * interface it replaces (paths relative to the SpecSeq++ tree).
 *
 * Index semantics: boundary indices and pair indices are the file's 0-based column
 * indices for both ASCII and binary input (the reference's internal +1 shift and dummy
 * column 0, main.cu:102-103,115, are not part of the contract).
 */
#ifndef SSQ_H
#define SSQ_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (replace exit(EXIT_FAILURE) in kernel.h:26-32 and the ignored read error
 * at main.cu:329). */
#define SSQ_OK 0
#define SSQ_EIO 1      /* file cannot be opened / read / written                       */
#define SSQ_EFORMAT 2  /* malformed PHAT file, index out of range, bad edge column      */
#define SSQ_ENOMEM 3   /* device or host memory exhausted (never silently degraded)   */
#define SSQ_EDEVICE 4  /* HIP runtime error                                            */
#define SSQ_EINVAL 5   /* invalid argument / parameter                                 */
#define SSQ_EBUSY 6    /* server queue full (retry later)                              */

/* Host CSC boundary matrix. Rows of each column are sorted, unique, in [0, n). */
typedef struct ssq_matrix {
  int64_t n;        /* number of columns (cells)            */
  int64_t nnz;      /* number of nonzeros                   */
  int64_t* col_ptr; /* n+1 offsets into rows                */
  int32_t* rows;    /* nnz row indices                      */
  int32_t* dims;    /* n cell dimensions                    */
} ssq_matrix;

/* The 10 numeric arguments of the specseq++ command line (main.cu:285-294, README.md:26).
 * level1/level2 are the slot capacities of the two-level column store; level 3 slots hold
 * 8*level2 entries (specseq.h:149-154). */
typedef struct ssq_params {
  int32_t dualize;       /* 0/1: reduce the anti-transposed matrix (cohomology)          */
  int32_t model;         /* bit0: high-dimension guided clearing, bit1: edge collapse    */
  int32_t max_iteration; /* neighbours examined per max-dim cell by guided clearing      */
  int32_t ss_block_size; /* first page width (columns and rows)                         */
  float growth;          /* page growth factor (>= 1)                                    */
  int32_t queue_size;    /* additions in flight per round (one wavefront each)          */
  int32_t level1;        /* level-1 slot capacity (int32 entries)                        */
  int32_t level2;        /* level-2 slot capacity (int32 entries)                        */
  /* Workgroup shape of the column additions: no reference counterpart (its shapes are
   * compile-time, bupt.h:48-71).  0 selects the engine's default; the pairs never depend on
   * them; the autotuner picks them per size class. */
  int32_t add_waves;       /* wavefronts per short-addition workgroup: 2, 4 or 8 (4)          */
  int32_t long_grid;       /* workgroups of the long-column kernel, <= 4096 (2 per CU)        */
  int32_t short_max;       /* largest |a|+|b| a wavefront merges in LDS, 64..1024 (1024)      */
  int32_t rounds_per_sync; /* add rounds between two reads of the worklist counts, 1..8 (3)   */
} ssq_params;

/* Persistence pairs, sorted by birth; optional essential (never-paired) cells. */
typedef struct ssq_pairs {
  int64_t np;
  int64_t* pairs; /* 2*np: birth, death */
  int64_t ne;
  int64_t* essential; /* ne cell indices, ascending */
} ssq_pairs;

/* Per-call statistics (the reference prints these; specseq.h:291-296,576-579,1053). */
typedef struct ssq_stats {
  double t_init_ms;      /* pre-passes: collapse, guided clearing, init passes          */
  double t_reduce_ms;    /* page loop over all dimensions                               */
  double t_total_ms;     /* whole call on the device side                               */
  double t_collapse_ms;  /* host part of the edge collapse (model bit 1)                */
  int64_t n_pending;     /* columns left for the page loop after the pre-passes        */
  int64_t n_apparent;    /* apparent pairs found by the init passes                    */
  int64_t n_cleared_hd;  /* max-dim cells cleared by guided clearing                   */
  int64_t n_kept_edges;  /* edges kept by the collapse                                 */
  int64_t n_pages;       /* pages run                                                  */
  int64_t n_rounds;      /* claim/add rounds run                                       */
  int64_t n_additions;   /* column additions performed                                 */
  double add_bytes;      /* algorithmic bytes of the short additions (LDS merges):
                            4(|a|+|b|+|a^b|)+24 each                                   */
  int64_t store_bytes;   /* device bytes reserved for the column store                 */
  double t_add_kernel_ms; /* summed duration of the column-addition kernel launches, from
                             HIP events on the engine stream (SSQ_KERNEL_TIMING=1), else 0 */
  int64_t n_add_launches; /* column-addition kernel launches                            */
  int64_t n_tiled_additions; /* additions merged tile by tile from HBM                   */
  int64_t n_dense_additions; /* additions into a workgroup's LDS bit-set accumulator      */
  double dense_bytes;     /* bytes those additions move: 4|b|+24 each (the pivot's rows; the
                             running column stays in the LDS bit set)                     */
  double t_add_long_kernel_ms; /* summed duration of the long-column kernel launches     */
  int64_t max_chain;      /* most additions one column made in one round                */
  int64_t max_addition_len; /* largest |a|+|b| of one addition                          */
  int64_t n_huge_additions; /* additions of columns whose rows spread past the LDS window
                               (an LDS window over a device-memory bit set)                */
  double huge_bytes;      /* bytes those additions move: 4|b|+24 each (the pivot's rows; the
                             window is in LDS, toggles below it are device atomics), and
                             4(|a|+|b|+|a xor b|)+24 for the ones merged tile by tile      */
  double t_add_huge_kernel_ms; /* summed duration of the huge-column kernel launches      */
} ssq_stats;

/* ---- I/O (main.cu:90-167 processFiles; main.cu:191-229 writeVectorToFile) ---------- */
int ssq_load_phat(const char* path, int binary, ssq_matrix* out);
int ssq_save_pairs(const char* path, int binary, const ssq_pairs* pairs);
void ssq_matrix_free(ssq_matrix* m);
void ssq_pairs_free(ssq_pairs* p);

/* Default parameters (README.md:44-53: 102400, 1.05, 65536, level1/level2 1000/2000). */
void ssq_default_params(ssq_params* p);

/* ---- Engine handle: one per host thread; owns a HIP stream and its device state. ---- */
typedef struct ssq_handle ssq_handle;
int ssq_create(int device, ssq_handle** out);
int ssq_destroy(ssq_handle* h);
const char* ssq_last_error(const ssq_handle* h);

/* Whole reduction of one host matrix: replaces the SpecSeqPlusPlus constructor
 * (specseq.h:486-512), GPU_Init (513-590), GPU_Compute (592-1060) and
 * GetPersistencePairs (1182-1193).  stats may be NULL. */
int ssq_reduce(ssq_handle* h, const ssq_matrix* m, const ssq_params* p, ssq_pairs* out,
               ssq_stats* stats);

/* The same reduction in the reference's two phases, for callers that time or schedule them apart
 * (main.cu:337-348 calls GPU_Init, then GPU_Compute).  ssq_init uploads the matrix and runs
 * specseq.h:513-590 GPU_Init's work: edge collapse (model bit 1), high-dim clearing (model bit 0)
 * or the anti-transpose (dualize), the init passes (initleft/initlookup/clearing/compress) and the
 * pending lists; stats->t_init_ms is its time.  ssq_compute runs specseq.h:592-1060 GPU_Compute
 * (the page loop) and returns the pairs as ssq_reduce does; stats->t_reduce_ms is its time.  A
 * handle holds one initialized matrix; any other reduction on the handle discards it, and
 * ssq_compute without one returns SSQ_EINVAL. */
int ssq_init(ssq_handle* h, const ssq_matrix* m, const ssq_params* p, ssq_stats* stats);
int ssq_compute(ssq_handle* h, ssq_pairs* out, ssq_stats* stats);

/* Same, with the matrix already resident in device memory (device pointers d_col_ptr
 * [n+1], d_rows[nnz], d_dims[n]).  Pairs are written to device memory d_pairs_out (2*n
 * int64 capacity) and their count to *np_out.  Used by the timed region of bench.py. */
int ssq_reduce_device(ssq_handle* h, int64_t n, int64_t nnz, const int64_t* d_col_ptr,
                      const int32_t* d_rows, const int32_t* d_dims, const ssq_params* p,
                      int64_t* d_pairs_out, int64_t* np_out, ssq_stats* stats);

/* ---- Device memory plumbing (no reference counterpart; the reference keeps its inputs in
 * managed memory, specseq.h:148-270).  For callers that keep matrices resident in HBM and
 * call ssq_reduce_device; all copies run on the handle's stream and are complete on return. */
int ssq_device_alloc(ssq_handle* h, size_t bytes, void** dptr);
int ssq_device_free(ssq_handle* h, void* dptr);
int ssq_copy_to_device(ssq_handle* h, void* dst, const void* src, size_t bytes);
int ssq_copy_to_host(ssq_handle* h, void* dst, const void* src, size_t bytes);
int ssq_device_synchronize(ssq_handle* h);

/* ---- Batching (new; no reference counterpart): many complexes, one engine pass. ---- */
/* The matrices are laid out block-diagonally on the device and reduced together; outs[i]
 * receives complex i's pairs in its own indices. */
int ssq_batch_reduce(ssq_handle* h, const ssq_matrix* const* mats, int count,
                     const ssq_params* p, ssq_pairs* outs, ssq_stats* stats);

/* ---- Matrices resident in HBM (new; the reference re-reads its input for every run,
 * main.cu:322-335).  ssq_matrix_to_device copies a host matrix into device memory of its own,
 * checked on the device (row range, strictly increasing rows); ssq_device_matrix_free releases
 * it (hipFree: it waits for the device).  ssq_batch_reduce_resident is ssq_batch_reduce on such
 * matrices: each is gathered into the batch by one device kernel, with no host staging and no
 * copy from the host; outs[i] holds matrix i's pairs in its own indices. */
typedef struct ssq_device_matrix {
  int64_t n, nnz;
  const int64_t* col_ptr; /* device memory, n+1 offsets, col_ptr[0] = 0 */
  const int32_t* rows;    /* device memory, nnz rows                    */
  const int32_t* dims;    /* device memory, n dimensions                */
} ssq_device_matrix;
int ssq_matrix_to_device(ssq_handle* h, const ssq_matrix* m, ssq_device_matrix* out);
int ssq_device_matrix_free(ssq_handle* h, ssq_device_matrix* m);
int ssq_batch_reduce_resident(ssq_handle* h, const ssq_device_matrix* mats, int count,
                              const ssq_params* p, ssq_pairs* outs, ssq_stats* stats);

/* ---- Serving (new): continuous batching request loop on one GPU. -------------------- */
typedef struct ssq_server ssq_server;
typedef struct ssq_server_config {
  int device;
  int streams;               /* batches that may be on the device at once (>=1)        */
  int max_batch;             /* requests per batch                                     */
  int64_t max_batch_cells;   /* cells per batch (0 = no limit): bounds batch latency   */
  int64_t mem_budget_bytes;  /* admission budget of estimated device bytes (0 = auto) */
  ssq_params params;         /* reduction of submitted matrices                        */
  ssq_params points_params;  /* reduction of point-cloud (Rips) requests               */
} ssq_server_config;
int ssq_server_create(const ssq_server_config* cfg, ssq_server** out);
/* Submit a matrix (copied) or a point cloud (npts x pdim float32, Rips threshold r,
 * max dimension maxdim).  *ticket receives the request id. */
int ssq_server_submit_matrix(ssq_server* s, const ssq_matrix* m, int64_t* ticket);
int ssq_server_submit_points(ssq_server* s, const float* pts, int64_t npts, int pdim,
                             float r, int maxdim, int64_t* ticket);
/* Matrices kept resident in HBM for requests that repeat: register checks the matrix on the
 * host once (offsets, rows strictly increasing and in range) and copies it into device memory
 * owned by the server; submit_registered queues a request on it (no host copy, no upload);
 * unregister waits for that matrix's outstanding requests and frees it (hipFree waits for the
 * device).  Registered requests use cfg.params and are batched with each other. */
int ssq_server_register_matrix(ssq_server* s, const ssq_matrix* m, int64_t* matrix_id);
int ssq_server_submit_registered(ssq_server* s, int64_t matrix_id, int64_t* ticket);
int ssq_server_unregister_matrix(ssq_server* s, int64_t matrix_id);
/* Non-blocking poll: returns SSQ_OK and fills out (caller frees) when ticket is done,
 * SSQ_EBUSY when still pending.  latency_ms (may be NULL): submit-to-completion. */
int ssq_server_poll(ssq_server* s, int64_t ticket, ssq_pairs* out, double* latency_ms);
int ssq_server_wait_all(ssq_server* s);
/* Counters since the server started, per batch kind (0: host matrices, 1: point clouds,
 * 2: registered matrices): batches, requests, cells, and the batches' service time (from
 * admission to published results) summed and at most. */
typedef struct ssq_server_stats {
  int64_t batches[3], requests[3], cells[3];
  double service_ms[3], max_service_ms[3];
} ssq_server_stats;
int ssq_server_get_stats(ssq_server* s, ssq_server_stats* out);
int ssq_server_destroy(ssq_server* s);

/* Host edge collapse (model bit 1): kept[i] = 1 for every edge that GUDHI's sequential
 * collapser keeps, edges given in filtration order (edge i has filtration i) by their two
 * vertex labels.  Replaces flag_complex_collapse_edges as called by specseq.h:410-424
 * (Flag_complex_edge_collapser.h:503-524).  threads <= 0: all host threads.  No GPU. */
int ssq_collapse_edges(int64_t ne, const int32_t* u, const int32_t* v, uint8_t* kept, int threads);

/* ---- Alpha complexes (replaces phat-highclear/addons/alpha_3.cpp, CGAL Delaunay_triangulation_3
 * with an exact kernel).  Host only.  npts x 3 float64 points; duplicates count once, vertex ids
 * follow each distinct point's first occurrence.  The Delaunay triangulation uses exact
 * predicates and symbolic perturbation (unique, independent of input order); values follow
 * alpha_3.cpp:260-352 (tetrahedra: squared circumradius; triangles and edges: the value of an
 * attached coface, else their own squared radius; vertices 0); cells with value <= max_value
 * (max_value < 0: all) in (value, dimension, vertex ids) order as alpha_3.cpp writes them
 * (:354-356), or with dim_major = 1 by dimension first, then value and vertex ids, as the
 * reference's data/Alpha files are stored.  values (may be NULL) receives each cell's value
 * (malloc'ed, n doubles; release with ssq_free).  SSQ_EINVAL when the points do not span 3-D
 * or their coordinates span more than 120 bits of one binary grid. */
int ssq_alpha_build(const double* pts, int64_t npts, double max_value, int dim_major, ssq_matrix* out,
                    double** values);
void ssq_free(void* p);
/* The Delaunay tetrahedra alone: 4 ascending vertex ids each (ids as above); *ntets receives
 * the count, SSQ_ENOMEM when it exceeds cap. */
int ssq_delaunay_3(const double* pts, int64_t npts, int32_t* tets, int64_t cap, int64_t* ntets);

/* ---- Inputs (phat-highclear/addons/rips.cpp:49-87 semantics) ------------------------- */
/* Vietoris-Rips boundary matrix of npts points in R^pdim: edges with float32 distance
 * <= r, cliques up to dimension maxdim, ordered by (filtration, dim, lexicographic vertex
 * ids) with filtration = longest edge. */
int ssq_vr_build(const float* pts, int64_t npts, int pdim, float r, int maxdim,
                 ssq_matrix* out);
/* The same matrix built on the handle's GPU (maxdim <= 2), copied back to the host. */
int ssq_vr_build_gpu(ssq_handle* h, const float* pts, int64_t npts, int pdim, float r, int maxdim,
                     ssq_matrix* out);
/* Persistence pairs of `count` point clouds' Rips complexes (same pdim, maxdim <= 2): the
 * complexes are built on the GPU directly as one block-diagonal batch in HBM and reduced in
 * one engine pass; nothing but points goes in and pairs come out.  outs[c] holds cloud c's
 * pairs in the indices of its own matrix (the one ssq_vr_build(pts[c], npts[c], pdim, r[c],
 * maxdim) returns).  No reference counterpart: the reference reduces one file per run
 * (main.cu:322-350); this serves the point-cloud requests of a batched stream. */
int ssq_batch_reduce_points(ssq_handle* h, int count, const float* const* pts, const int64_t* npts, int pdim,
                            const float* r, int maxdim, const ssq_params* p, ssq_pairs* outs,
                            ssq_stats* stats);

/* ---- Batched Rips cohomology with everything in HBM (no reference counterpart; the batched
 * stream's hot path).  Clouds of at most 4096 points each, maxdim 2.  d_pts holds every cloud's
 * points (pbase[count] x pdim float32, device memory); cloud c is points [pbase[c], pbase[c+1])
 * with threshold radius[c] (pbase, radius: host arrays).  The coboundary matrices are built on
 * the GPU straight into the engine's dual layout, in groups sized to the free device memory,
 * and reduced in cohomology (the pairs equal those of ssq_reduce on ssq_vr_build's matrix;
 * dualize is implied, model bit 1 is not supported).  Pairs, sorted by birth in each cloud's
 * own indices, go to device memory: cloud c's are d_pairs[2*pair_off[c] .. 2*pair_off[c+1]).
 * pairs_cap must be at least the number of vertices plus edges (ssq_rips_count). */
int ssq_rips_count(ssq_handle* h, int count, const float* d_pts, const int64_t* pbase, int pdim,
                   const float* radius, int64_t* edges, int64_t* triangles);
int ssq_batch_reduce_points_device(ssq_handle* h, int count, const float* d_pts, const int64_t* pbase, int pdim,
                                   const float* radius, int maxdim, const ssq_params* p, int64_t* d_pairs,
                                   int64_t pairs_cap, int64_t* pair_off, ssq_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
