/*
 * cugal_hip.h - C ABI of libcugal_hip.so, the MI355X (gfx950) kernels of the
 * FUGAL graph-alignment pipeline.
 *
 * Conventions (all entry points):
 *   - plain device pointers and sizes; no framework types;
 *   - matrices are n x n FP32, row-major, contiguous (leading dimension n);
 *   - CSR adjacency: int32 col_indices[nnz], int32 row_pointers[n + 1], the
 *     columns of a row sorted ascending;
 *   - node-feature matrices are n x 4 FP32 row-major
 *     (degree, clustering, mean neighbour degree, mean neighbour clustering),
 *     or n x 8 where an entry point takes a feature_dim (FUGAL's 7 features
 *     and a zero, cg_fugal_features);
 *   - `stream` is a hipStream_t (NULL = the null stream); work is enqueued on
 *     it and nothing synchronises the device unless the entry point says so;
 *   - temporaries come from a caller-provided device workspace whose size the
 *     matching *_workspace() function returns; the library never allocates
 *     device memory;
 *   - the return value is 0 on success or a negative code, with a message
 *     available from cg_last_error() (thread-local);
 *   - calls from several host threads may run at once when each uses its
 *     own workspace (and stream): host-side state such as the Sinkhorn
 *     solves' pinned poll slots is per thread.
 *
 * Each declaration names the reference interface it replaces
 * (SebastianRueClausen/cuGAL, path:line).
 */
#ifndef CUGAL_HIP_H
#define CUGAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* cg_stream_t;

/* Message of the last failed call on this thread ("" if none). */
const char* cg_last_error(void);
/* ABI version of the library (major * 100 + minor). */
int cg_version(void);

/* ------------------------------------------------------------------------
 * Adjacency.
 * ---------------------------------------------------------------------- */

/* Replaces create_adjacency (cuda_kernels/adjacency.cu:228-259, bound at
 * cuda_kernels/kernels.cpp:38-40) and the CSR part of Adjacency.from_graph
 * (cugal/adjacency.py:46-69).  `edges` holds edge_count int32 pairs (x, y);
 * each pair contributes the entry A[y][x] and, when `symmetrize` is set, also
 * A[x][y] (the reference appends the reversed list in Python,
 * cugal/adjacency.py:53-55).  Writes row_pointers[node_count + 1] and
 * col_indices[nnz] with nnz = edge_count * (symmetrize ? 2 : 1); columns are
 * sorted within each row, and rows without entries (including leading and
 * trailing isolated vertices) get empty ranges. */
size_t cg_csr_build_workspace(int64_t edge_count, int64_t node_count, int symmetrize);
int cg_csr_build(const int32_t* edges, int64_t edge_count, int64_t node_count, int symmetrize,
                 int32_t* col_indices, int32_t* row_pointers, void* workspace, size_t workspace_bytes,
                 cg_stream_t stream);

/* Replaces adjacency_matmul (cuda_kernels/adjacency.cu:77-150, bound at
 * cuda_kernels/kernels.cpp:34-37; used by Adjacency.mul,
 * cugal/adjacency.py:128-143): out = (negate ? -A : A) @ matrix for a
 * contiguous n x n matrix. */
int cg_adjacency_matmul(const int32_t* col_indices, const int32_t* row_pointers, int64_t n,
                        const float* matrix, float* out, int negate, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Node features.
 * ---------------------------------------------------------------------- */

/* Replaces graph_features (cuda_kernels/feature_extraction.cu:67-84, bound
 * at cuda_kernels/kernels.cpp:43): per vertex the degree and the clustering
 * coefficient t / (d (d - 1)) where t = sum over neighbours u of
 * |N(v) & N(u)| (twice the triangle count).  `triangles` (nullable)
 * receives t as an exact int32. */
int cg_graph_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float* clustering,
                      float* degrees, int32_t* triangles, cg_stream_t stream);

/* Replaces average_neighbor_features (cuda_kernels/feature_extraction.cu:
 * 108-125, bound at cuda_kernels/kernels.cpp:44-47): averages[v] = mean of
 * features over N(v), 0 for isolated vertices. */
int cg_average_neighbor_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n,
                                 const float* features, float* averages, cg_stream_t stream);

/* Replaces extract_features_cuda + the mu scaling of Features.create
 * (cugal/feature_extraction.py:15-38, :91-92) in one call: out (n x 4) =
 * mu * (degree, clustering, mean neighbour degree, mean neighbour
 * clustering).  `triangles` (nullable) as in cg_graph_features.  Integer
 * sums are exact; the divisions are correctly rounded FP32. */
size_t cg_node_features_workspace(int64_t n);
int cg_node_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float mu, float* out,
                     int32_t* triangles, void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Per CSR entry (v, u) the number of common neighbours |N(v) & N(u)| (out:
 * int32, one per entry of col_indices; rows sorted).  No reference
 * counterpart: the locality order of the gradient's gathers drops the edges
 * that close no triangle (the shortcuts of a small-world graph) before its
 * breadth-first numbering. */
int cg_edge_common_neighbours(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, int32_t* out,
                              cg_stream_t stream);

/* Ego-net edge count per vertex, deg(v) + t(v) / 2: the feature of
 * FUGAL/pred.py:68-72 (nx.ego_graph(G, v).number_of_edges() on a simple
 * graph).  Needs `triangles` from cg_graph_features / cg_node_features. */
int cg_egonet_edges(const int32_t* row_pointers, const int32_t* triangles, int64_t n, int32_t* out,
                    cg_stream_t stream);

/* The 7-feature node set of the original FUGAL (feature_extraction with
 * simple=False, FUGAL/pred.py:24-110) for a simple undirected graph, as
 * 8-float rows scaled by mu (FUGAL/pred.py:134 scales the distance by mu,
 * which is the same): degree, clustering, mean neighbour degree, mean
 * neighbour clustering (as cg_node_features), ego-net edges
 * (nx.ego_graph(G, v).number_of_edges()), edges leaving the ego net, distinct
 * vertices at distance 2, and 0.  Isolated vertices get zeros.  out is n x 8
 * FP32 (16-byte aligned); n is limited by two n-bit LDS bitmaps (n <= 655360). */
size_t cg_fugal_features_workspace(int64_t n);
int cg_fugal_features(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, float mu, float* out,
                      void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Cost / gradient.
 * ---------------------------------------------------------------------- */

/* Replaces add_distance (cuda_kernels/distance.cu:35-53, bound at
 * cuda_kernels/kernels.cpp:60): out[x][y] += ||source[x] - target[y]||_2
 * over the 4 features. */
int cg_add_distance(const float* source, const float* target, int64_t n, float* out, cg_stream_t stream);

/* Replaces regularize (cuda_kernels/regularize.cu:24-39, bound at
 * cuda_kernels/kernels.cpp:63): gradient += iteration - 2 * iteration * P,
 * 64-bit indexing. */
int cg_regularize(float* gradient, const float* P, int64_t n, int iteration, cg_stream_t stream);

/* The whole Frank-Wolfe gradient of cugal/pred.py:38-74 in two kernels for
 * undirected graphs (A, B symmetric, given as CSR):
 *   gradient = -2 A P B + D + regularizer_factor * lambda * (1 - 2 P)
 * where D[x][y] = ||source_features[x] - target_features[y]||_2 is recomputed
 * on the fly (no n x n distance matrix; features may be NULL for D = 0);
 * feature rows hold feature_dim = 4 or 8 floats.
 * regularizer_factor = 1 is sparse_gradient (cugal/pred.py:54-74); 2 is
 * dense_gradient, which adds lambda (1 - 2P) twice (cugal/pred.py:46-50).
 * ld_p: row stride of P in floats (>= n; n for a contiguous P).  The first
 * product gathers 128-byte segments of P's rows; with a stride that is a
 * multiple of 4 KiB those alias a few L2 sets, so such a P is first copied
 * to a row-padded buffer, while any other stride (cg_gradient_padded_ld)
 * is gathered in place.  Each product serves the neighbours near a 64-row
 * tile from rows staged in LDS (exact for any CSR; a row whose sorted run is
 * not local is gathered from global memory), except for a graph with a row
 * of more than 64 entries, which takes the plain gather: the first call with
 * a given CSR (by its device pointers and n) synchronises the stream once to
 * read that property, later calls do not.  Workspace: the padded copy and
 * the intermediate (A P)^T, n x (n + 64) FP32 each, and 8 n bytes per graph
 * for the local runs. */
size_t cg_gradient_workspace(int64_t n);
int64_t cg_gradient_padded_ld(int64_t n);
int cg_gradient(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* b_col, const int32_t* b_rowptr,
                int64_t n, const float* P, int64_t ld_p, const float* source_features, const float* target_features,
                int feature_dim, int lambda_, float regularizer_factor, float* gradient, void* workspace,
                size_t workspace_bytes, cg_stream_t stream);

/* The gradient of cugal/pred.py:38-51 for directed graphs (any A, B, as
 * nx.to_numpy_array gives them, cugal/pred.py:106-110), and of the general
 * four-product branch of sparse_gradient (cugal/pred.py:65-67):
 *   gradient = -(A P B^T + A^T P B) + D + regularizer_factor * lambda (1 - 2P)
 * with A, A^T, B, B^T as CSR (row i lists the columns j with M[i][j] != 0,
 * repeated entries counted).  The two products share their second pass
 * (one gather of B's and B^T's rows per output tile).  Workspace: the padded
 * copy of P and the two intermediates (A P)^T, (A^T P)^T, n x (n + 64) FP32
 * each. */
size_t cg_gradient_directed_workspace(int64_t n);
int cg_gradient_directed(const int32_t* a_col, const int32_t* a_rowptr, const int32_t* at_col,
                         const int32_t* at_rowptr, const int32_t* b_col, const int32_t* b_rowptr,
                         const int32_t* bt_col, const int32_t* bt_rowptr, int64_t n, const float* P, int64_t ld_p,
                         const float* source_features, const float* target_features, int feature_dim, int lambda_,
                         float regularizer_factor, float* gradient, void* workspace, size_t workspace_bytes,
                         cg_stream_t stream);

/* dst[i][j] = src[row_map[i]][col_map[j]] for n x n FP32 matrices with row
 * strides ld_src, ld_dst (a NULL map is the identity; maps hold values in
 * 0..n-1, which the caller guarantees).  No reference counterpart: the
 * Frank-Wolfe loop may run on relabelled graphs (a locality order of their
 * CSR rows, cugal/pred.py:89-148 is unchanged by a relabelling up to the
 * order of floating-point sums), and this copies its quasi-permutation back
 * to the callers' labels. */
int cg_permute_copy(const float* src, int64_t ld_src, int64_t n, const int32_t* row_map, const int32_t* col_map,
                    float* dst, int64_t ld_dst, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Sinkhorn (log domain, FP32).
 * ---------------------------------------------------------------------- */

/* Replaces sinkhorn_log_step (cuda_kernels/sinkhorn_log.cu:100-114, bound at
 * cuda_kernels/kernels.cpp:50-53): for each row r of the n x n matrix K,
 *   out[r] = (1 - momentum) out[r] - momentum * LSE_c(K[r][c] + add[c]). */
int cg_sinkhorn_log_step(const float* K, int64_t n, const float* add, float* out, float momentum,
                         cg_stream_t stream);

/* Replaces SinkhornState.solve_standard (cugal/sinkhorn.py:122-144) in the
 * reference's default FP64 (cugal/config.py:65): C, u, v FP64; K =
 * exp(-C / reg) evaluated where C is read; per iteration v = v^(1-m)
 * (1 / (u^T K + 1e-16))^m, u = u^(1-m) (1 / (K v + 1e-16))^m, the
 * column-marginal error sum_j |sum_i u_i K_ij v_j - 1| at iterations 0, f,
 * 2f, ... and the stop below `threshold`; Lehmann momentum from
 * `momentum_start` (-1: off), all decided on the device.  Arguments and
 * outputs as cg_sinkhorn_standard_solve. */
size_t cg_sinkhorn_standard_f64_workspace(int64_t n, int max_iterations, int eval_freq);
int cg_sinkhorn_standard_solve_f64(const double* C, int64_t n, double regularization, double* u, double* v,
                                   int max_iterations, double threshold, int eval_freq, int momentum_start,
                                   double* errors, int errors_capacity, int* error_count, int* iterations,
                                   void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Replaces sinkhorn_log_marginal (cuda_kernels/sinkhorn_log.cu:88-98, bound
 * at cuda_kernels/kernels.cpp:54-57): sum_r |sum_c exp(K[r][c] + v[c] + u[r])
 * - 1| accumulated in FP64; synchronises the stream and writes *result on the
 * host. */
size_t cg_sinkhorn_log_marginal_workspace(int64_t n);
int cg_sinkhorn_log_marginal(const float* K, int64_t n, const float* u, const float* v, double* result,
                             void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Replaces SinkhornState.solve_log (cugal/sinkhorn.py:90-120) including its
 * per-iteration kernel launches.  C is the cost (the FW gradient); K = -C /
 * regularization is never materialised.  u, v (device, n floats, natural-log
 * potentials) are the warm start and receive the result.  Per iteration
 * v = -LSE_i(K + u), u = -LSE_j(K + v); every eval_freq iterations (not at
 * iteration 0) the column-marginal L1 error is recorded (FP64) and the solve
 * stops when it is below `threshold`.  momentum_start < 0 disables Lehmann
 * momentum (cugal/sinkhorn.py:78-88); otherwise it is the reference's
 * sinkhorn_momentum_start.  fused = 1 reads C once per iteration (4 <= n <=
 * 32768, any width, 16-byte aligned C; otherwise the two-sweep structure is
 * used, same results); fused = 0 reads it twice (column then row sweep).
 * Synchronises the stream once at the end.  Host outputs: errors[0 ..
 * min(count, errors_capacity)), *error_count, *iterations (the reference's
 * SinkhornProfile.iteration_count), and when non-NULL the summed duration and
 * number of the C-reading sweep kernels of every 10th iteration (iterations
 * 5, 15, 25, ...; for roofline reporting: event records around those
 * launches only, so the timer costs ~1/10 of timing every sweep). */
size_t cg_sinkhorn_workspace(int64_t n, int fused, int max_iterations, int eval_freq);
int cg_sinkhorn_fused_supported(int64_t n);
int cg_sinkhorn_log_solve(const float* C, int64_t n, float regularization, float* u, float* v, int max_iterations,
                          double threshold, int eval_freq, int momentum_start, int fused, double* errors,
                          int errors_capacity, int* error_count, int* iterations, float* sweep_ms,
                          int* sweep_launches, void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Replaces solve_standard (cugal/sinkhorn.py:122-144): with K = exp(-C/reg)
 * (recomputed from C where read, never stored) and u, v the scalings,
 *   v_j = v_j^(1-m) (1 / (sum_i u_i K_ij + 1e-16))^m,
 *   u_i = u_i^(1-m) (1 / (sum_j K_ij v_j + 1e-16))^m,
 * the column-marginal error sum_j |sum_i u_i K_ij v_j - 1| (FP64) at
 * iterations 0, eval_freq, 2 eval_freq, ..., and a stop below `threshold`;
 * Lehmann momentum as in cg_sinkhorn_log_solve.  u, v (device, n floats)
 * are read as the warm start and overwritten.  The fused path reads C once
 * per iteration; otherwise a row and a column sweep.  Workspace:
 * cg_sinkhorn_workspace(n, fused, max_iterations, eval_freq).  FP32 as the
 * reference's GPU run of this method (overflow/underflow of K as there). */
int cg_sinkhorn_standard_solve(const float* C, int64_t n, float regularization, float* u, float* v,
                               int max_iterations, double threshold, int eval_freq, int momentum_start, int fused,
                               double* errors, int errors_capacity, int* error_count, int* iterations,
                               void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Differentiable unrolled log-Sinkhorn.  No reference counterpart: the
 * reference's only differentiable form is torch.logsumexp in its CPU branch
 * (cugal/sinkhorn.py:107-111), whose iteration this adjoint follows.
 * ---------------------------------------------------------------------- */

/* C (n x n) = lam * sum_k weights[k] |source[i][k] - target[j][k]| (weighted
 * L1 over the 4 features); weights is a device array of 4 floats. */
int cg_feature_cost(const float* source_features, const float* target_features, const float* weights, float lam,
                    int64_t n, float* cost, cg_stream_t stream);

/* plan = exp(-C / reg + u_i + v_j) (natural-log duals). Workspace >= 8n + 512 bytes. */
int cg_sinkhorn_plan(const float* C, int64_t n, float regularization, const float* u, const float* v, float* plan,
                     void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Forward: `iterations` plain log-domain iterations (v then u, no stopping
 * test) from the duals u0, v0; writes the final duals u, v and opaque
 * checkpoints (cg_sinkhorn_checkpoint_floats() floats) holding the duals
 * every `checkpoint_every` iterations (1..64).  Backward: given
 * grad_plan = dL/dplan (plan = exp(-C/reg + u + v) of the final duals),
 * recomputes each checkpoint segment and writes dL/dC into grad_cost
 * (nullable) and, when source_features is given (C built by
 * cg_feature_cost with `weights`, `lam`), grad_lam_w[0] = dL/dlam and
 * grad_lam_w[1..4] = dL/dweights (device doubles). */
size_t cg_sinkhorn_checkpoint_floats(int64_t n, int iterations, int checkpoint_every);
size_t cg_sinkhorn_unrolled_workspace(int64_t n, int checkpoint_every);
int cg_sinkhorn_unrolled_forward(const float* C, int64_t n, float regularization, int iterations,
                                 int checkpoint_every, const float* u0, const float* v0, float* checkpoints,
                                 float* u, float* v, void* workspace, size_t workspace_bytes, cg_stream_t stream);
int cg_sinkhorn_unrolled_backward(const float* C, int64_t n, float regularization, int iterations,
                                  int checkpoint_every, const float* checkpoints, const float* grad_plan,
                                  float* grad_cost, const float* source_features, const float* target_features,
                                  const float* weights, float lam, double* grad_lam_w, void* workspace,
                                  size_t workspace_bytes, cg_stream_t stream);

/* The precision form of cg_gradient (same quantity, cugal/pred.py:38-74):
 * G = -2 A P B + D + regularizer_factor * lambda_ * (1 - 2P) with A, B as
 * exact 0/1 BF16 operands and P split into BF16 hi + lo, products on the
 * matrix cores (v_mfma_f32_32x32x16_bf16) accumulated in FP32.  Where the
 * reference multiplies dense adjacencies in FP32 (torch.matmul,
 * cugal/pred.py:45-46) this runs on MFMA; each product carries a relative
 * error of about 2^-16 from the hi+lo split.  a_bits, b_bits:
 * cg_adjacency_bits_words(n) words each, from cg_adjacency_pack_bits: the
 * adjacency as bits, zero padded to np = n rounded up to 512; word
 * (k / 32) * np + i holds bits k % 32 of row i.  Undirected graphs only. */
size_t cg_adjacency_bits_words(int64_t n);
int cg_adjacency_pack_bits(const int32_t* col_indices, const int32_t* row_pointers, int64_t n, uint32_t* out,
                           cg_stream_t stream);
size_t cg_gradient_mfma_workspace(int64_t n);
int cg_gradient_mfma(const uint32_t* a_bits, const uint32_t* b_bits, int64_t n, const float* P, int64_t ld_p,
                     const float* source_features, const float* target_features, int feature_dim, int lambda_,
                     float regularizer_factor, float* gradient, void* workspace, size_t workspace_bytes,
                     cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Frank-Wolfe update.
 * ---------------------------------------------------------------------- */

/* Replaces update_quasi_permutation_log (cuda_kernels/update_quasi_permutation
 * .cu:40-56, bound at cuda_kernels/kernels.cpp:66; called at
 * cugal/pred.py:134-136): with plan = exp(-C/reg + u_i + v_j),
 *   gap = sum (plan - P) * C  (FP64),   P += alpha * (plan - P).
 * *gap_device (a device double) receives the signed sum; the reference
 * reports |gap|.  Takes the cost C instead of K = -C/reg (contiguous); P has
 * row stride ld_p >= n. */
size_t cg_update_workspace(int64_t n);
int cg_update_quasi_permutation_log(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                    const float* v, float alpha, float regularization, double* gap_device,
                                    void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* No counterpart in the reference: cg_update_quasi_permutation_log for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that forbid hipMalloc).  It takes no workspace and allocates nothing: its
 * partial sums live in a fixed device array of the library (16.25 KiB per
 * device, held until the process ends; n up to 2^30, larger n is rejected).
 * Every call on one device shares that array, so calls that overlap on one
 * device (concurrent streams or threads) race; calls in stream order do not. */
int cg_update_quasi_permutation_log_static(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                           const float* v, float alpha, float regularization, double* gap_device,
                                           cg_stream_t stream);

/* The STANDARD form of the same step (cugal/pred.py:77-86 with
 * scale_kernel_matrix, cugal/sinkhorn.py:20-23; gap as :139-141):
 * plan = u_i exp(-C/reg) v_j with the scalings of cg_sinkhorn_standard_solve. */
int cg_update_quasi_permutation_standard(float* P, int64_t ld_p, const float* C, int64_t n, const float* u,
                                         const float* v, float alpha, float regularization, double* gap_device,
                                         void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* The update in FP64 (the reference's default dtype, cugal/config.py:65):
 * standard != 0 takes the scalings of cg_sinkhorn_standard_solve_f64
 * (plan = u_i exp(-C/reg) v_j), else FP64 log duals (plan = exp(-C/reg +
 * u_i + v_j)); P, C, u, v FP64, the same workspace as the FP32 forms. */
int cg_update_quasi_permutation_f64(int standard, double* P, int64_t ld_p, const double* C, int64_t n,
                                    const double* u, const double* v, double alpha, double regularization,
                                    double* gap_device, void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Rounding to a permutation (LAP).
 * ---------------------------------------------------------------------- */

/* Replaces greedy_lap (cugal/greedy_lap.py:7-85) for
 *   method 0 = GREEDY (rows in index order, :10-19),
 *   method 1 = DOUBLE_GREEDY (rows by descending row maximum, :50-63; ties
 *              in row order, i.e. a stable sort),
 *   method 2 = PARALLEL_GREEDY (:65-85: rows whose maximum is > 0.5 take the
 *              first index of it directly and are not masked; the others are
 *              visited by descending maximum with those columns taken).
 * For each row in order, the columns already taken are multiplied by 0 in P
 * (in place, as the reference does) and the first index of the row maximum
 * is taken.  assignment (device, n int32) receives the column of each row. */
size_t cg_lap_workspace(int64_t n);
int cg_lap_greedy(float* P, int64_t n, int method, int32_t* assignment, void* workspace, size_t workspace_bytes,
                  cg_stream_t stream);

/* Replaces the RAND (cugal/greedy_lap.py:21-32) and MORE_RAND (:34-48)
 * branches: method 0 = RAND, 1 = MORE_RAND.  order (device, n int32) is the
 * row order the reference draws with torch.randperm; it must be a
 * permutation of 0..n-1 (checked, with a stream synchronisation).  MORE_RAND
 * also takes uniforms (device, n doubles in [0, 1]): the k-th visited row's
 * random.random() draw U_k, so that r_k = total_k * U_k stands for the
 * reference's uniform(0, cumsum[-1]); the column is the first j whose FP64
 * prefix sum of the masked row is >= r_k (-1 if none, which needs negative
 * entries or a NaN).  Same workspace and in-place masking as cg_lap_greedy. */
int cg_lap_greedy_random(float* P, int64_t n, int method, const int32_t* order, const double* uniforms,
                         int32_t* assignment, void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* Replaces dense_hungarian (cuda_kernels/hungarian/hungarian.cu:436-448,
 * bound at cuda_kernels/kernels.cpp:69) and the cugraph dense Hungarian of
 * HungarianMethod.DENSE (cugal/pred.py:161-166): an exact linear assignment
 * of the n x n FP32 matrix `cost` (row-major, not modified), minimising the
 * sum (maximize = 0) or maximising it (maximize = 1; the reference passes
 * 1 - P to a minimiser, which rounds small entries of P away in FP32; here
 * P is maximised directly).  Shortest augmenting paths with FP64 duals,
 * warm-started from column and row reductions.  assignment (device, n
 * int32) receives the column of each row; stats (device, 2 int32, may be
 * null) receives the number of rows left free by the warm start and the
 * total Dijkstra steps.  n <= 32768.  Synchronises the stream (it reports
 * NaN or infinite costs as CG_EINVAL). */
size_t cg_lap_hungarian_workspace(int64_t n);
int cg_lap_hungarian(const float* cost, int64_t n, int maximize, int32_t* assignment, int32_t* stats,
                     void* workspace, size_t workspace_bytes, cg_stream_t stream);

/* ------------------------------------------------------------------------
 * Alignment quality (benchmark harness).
 * ---------------------------------------------------------------------- */

/* Replaces the dense numpy metrics edge_correctness,
 * induced_conserved_structure and symmetric_substructure
 * (benchmark/experiment.py:115-141, called from Result.calculate, :365-376).
 * S (n_source rows) and T (n_target rows) are CSRs in the graphs' node order
 * (the row order nx.to_numpy_array gives them); answer[m] is the computed
 * alignment and truth[m] the ground-truth permutation (device int32).
 * counts (device, 4 int64) receives
 *   [0] inter  = #{(a, b): S[answer[a]][answer[b]] = 1 and T[truth[a]][truth[b]] = 1},
 *   [1] ones_s = #{S == 1},  [2] ones_t = #{T[truth][:, truth] == 1},
 *   [3] flags: bit 0 an answer outside [-n_source, n_source) (the reference
 *       then scores 0), bit 1 truth is not a permutation of distinct nodes;
 * negative answers index from the end, as numpy does.  EC = inter / ones_s,
 * ICS = inter / ones_t, S3 = inter / (ones_s + ones_t - inter). */
size_t cg_alignment_counts_workspace(int64_t n_target, int64_t m);
int cg_alignment_counts(const int32_t* s_col, const int32_t* s_rowptr, int64_t n_source, const int32_t* t_col,
                        const int32_t* t_rowptr, int64_t n_target, const int32_t* answer, const int32_t* truth,
                        int64_t m, int64_t* counts, void* workspace, size_t workspace_bytes, cg_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUGAL_HIP_H */
