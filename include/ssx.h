/*
 * ssx.h -- C ABI of libssx.so, the gfx950 (MI355X) HIP implementation of the slope-stability hot
 * path: Mohr-Coulomb return mapping, fixed-pattern tangent assembly, CSR operator and the vector
 * kernels of the Krylov / multigrid solvers.
 *
 * Conventions for every entry point
 *   - pointers are device pointers (HBM), owned by the caller; the library never allocates or
 *     frees caller data;
 *   - sizes are int64_t; `stream` is a hipStream_t (NULL = default stream); all work is
 *     stream-ordered, nothing synchronises unless stated;
 *   - return 0 on success, 1 for invalid arguments (message via ssx_last_error()), 2 when a
 *     kernel launch failed, 3 for other runtime errors;
 *   - point-major layouts: E[p*ns + c], S[p*ns + c], DS[p*ns*ns + i + ns*j] = dS_i/dE_j;
 *     dphi[(d*n_int + p)*n_p + node]; 3D strain [e11 e22 e33 g12 g23 g13], 2D [e11 e22 g12],
 *     2D stress [s11 s22 s12 s33]; global dof = dim*node + comp.
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   cython/constitutive_3d_batch.h:4-25, cython/assemble_tangent_values_3d.h:6-68 (declared to
 *   Python at cython/_kernels.pyx:14-97) and the NumPy/PETSc calls named per entry point below.
 */
#ifndef SSX_H
#define SSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library ABI version. */
int ssx_version(void);
/* Message of the last failed call on this host thread. */
const char *ssx_last_error(void);
/* hipStreamSynchronize(stream). */
int ssx_stream_sync(void *stream);

/* Davis strength reduction per point: (c0, phi, psi [rad], lambda) -> (c_bar, sin_phi).
 * davis_type 0/1/2 = A/B/C.  Replaces constitutive/reduction.py:14-78 (`reduction`). */
int ssx_reduction(const double *c0, const double *phi, const double *psi, int64_t n, double lam, int davis_type,
                  double *c_bar, double *sin_phi, void *stream);

/* 3D Mohr-Coulomb return mapping: E (n x 6) -> S (n x 6), DS (n x 36, may be NULL), optional FP32
 * copy of DS and per-point branch id (0 elastic, 1 smooth, 2 left edge, 3 right edge, 4 apex).
 * Replaces constitutive_problem_3d_s_batch_c / constitutive_problem_3d_sds_batch_c
 * (cython/constitutive_3d_batch.h:4-25, cython/constitutive_3d_batch.c:9-50). */
int ssx_constitutive_3d(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                        const double *lame, int64_t n, double *S, double *DS, float *DS32, int8_t *branch,
                        void *stream);

/* Same with material data as a per-element id into a table of rows (c_bar, sin_phi, shear, lame);
 * n = n_elem * n_q.  Same reference entry points as ssx_constitutive_3d, with the per-point
 * material arrays of cython/_kernels.pyx:100-123 folded into a table. */
int ssx_constitutive_3d_mat(const double *E, const int32_t *mat_id, const double *table, int n_mat, int n_q,
                            int64_t n, double *S, double *DS, float *DS32, int8_t *branch, void *stream);

/* ssx_constitutive_3d with the tangent written packed: DSp (n x 21) holds DS[s + 6 m] for s <= m,
 * row by row (s = 0: m = 0..5; s = 1: m = 1..5; ...): the full layout's upper entries to within
 * the last bit (only the 21 entries are formed, so a product the full kernel shares between (i, j)
 * and (j, i) may fuse into an FMA here; at most 2.6e-16 of the point's largest entry), with the
 * same stresses and branch ids bit for bit.  The associated-flow tangent is symmetric (sin psi = sin phi in the reference kernel,
 * cython/constitutive_3D_kernel.h:73), so the 21 entries carry it whole; no counterpart in the
 * reference, for callers that hand DS straight to ssx_element_blocks_affine_3d (ds_sym = 1). */
int ssx_constitutive_3d_sym(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                            const double *lame, int64_t n, double *S, double *DSp, int8_t *branch, void *stream);

/* The full point-major DS (n x 36) of a packed one: DS[s + 6 m] = DSp[(min, max) of (s, m)]. */
int ssx_ds_unpack_sym(int64_t n, const double *DSp, double *DS, void *stream);

/* Plane-strain return mapping: E (n x 3) -> S (n x 4), DS (n x 9).
 * Replaces constitutive_problem_2D (constitutive/problem.py:143-475). */
int ssx_constitutive_2d(const double *E, const double *c_bar, const double *sin_phi, const double *shear,
                        const double *bulk, const double *lame, int64_t n, double *S, double *DS, float *DS32,
                        int8_t *branch, void *stream);

/* Strain gather E = B u for (dim, n_p) in {(3,4),(3,10),(3,35),(2,3),(2,6),(2,15)}; edofs is
 * (n_elem x dim*n_p) int32 dof numbers into u.
 * Replaces assemble_overlap_strain_3d_c (cython/assemble_tangent_values_3d.h:50-61) and the 2D
 * `overlap_B @ u` (fem/distributed_tangent.py:906-907). */
int ssx_strain(int dim, int n_p, const double *dphi, const int32_t *edofs, const double *u, int64_t n_elem, int n_q,
               double *E, void *stream);

/* Internal force, stage 1: element vectors fe (n_elem x dim*n_p) = sum_q w B^T S.
 * Together with ssx_force_gather replaces assemble_force_3d_rows_c
 * (cython/assemble_tangent_values_3d.h:63-68) and the 2D B^T(wS) (constitutive/problem.py:1954). */
int ssx_force_elem(int dim, int n_p, const double *dphi, const double *S, const double *w, int64_t n_elem, int n_q,
                   double *fe, void *stream);

/* Internal force, stage 2: F[row] = sum over the row's slots of fe[slot_elem*nld + slot_lrow];
 * rows without slots (constrained) get 0 (fem/distributed_tangent.py:932). */
int ssx_force_gather(const double *fe, const int32_t *row_slot_ptr, const int32_t *slot_elem, const uint8_t *slot_lrow,
                     int64_t n_rows, int nld, double *F, void *stream);
/* Internal force, stage 2 through the node slots (ns_ptr / ns_elem / ns_lnode, element order per
 * node): the same sums as ssx_force_gather, bit for bit, with the dim rows of a node read together;
 * rows with row_free == 0 get 0 (fem/distributed_tangent.py:932). */
int ssx_force_gather_nodes(int dim, int nld, const double *fe, const int32_t *ns_ptr, const int32_t *ns_elem,
                           const uint8_t *ns_lnode, const uint8_t *row_free, int64_t n_nodes, double *F, void *stream);

/* Tangent values on the fixed CSR pattern through the row-slot lists; constrained rows get a unit
 * diagonal.  With `elastic` non-NULL writes K_r = (1-r) K + r K_el; `out32` (may be NULL) receives
 * an FP32 copy.  max_block_nnz = max over 64-row blocks of the block's nnz (sizes the LDS tile).
 * Replaces assemble_tangent_values_3d_rows_c (cython/assemble_tangent_values_3d.h:21-38) with the
 * diagonal fix-up of fem/distributed_tangent.py:1606-1607 and the K_r blend of
 * constitutive/problem.py:1614-1615. */
int ssx_tangent_rows(int dim, int n_p, const double *dphi, const double *DS, const double *w,
                     const int32_t *row_slot_ptr, const int32_t *slot_elem, const uint8_t *slot_lrow,
                     const int32_t *slot_pos, const int32_t *indptr, const uint8_t *row_free, int64_t n_rows,
                     int64_t n_elem, int n_q, double r, const double *elastic, double *out, float *out32,
                     int64_t max_block_nnz, void *stream);

/* y = alpha*A x + beta*y for CSR A (int32 indptr/indices); lanes per row: power of two <= 64.  With
 * at least 4 lanes and 16-byte aligned indices/val, each lane reads aligned quadruples of entries
 * (the same row form serves the fused Chebyshev step); otherwise one entry per lane and load.
 * Replaces PETSc MatMult (linear/solver.py:3431-3451). */
int ssx_spmv_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, const double *x,
                 double *y, double alpha, double beta, int lanes, void *stream);
int ssx_spmv_f32(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const float *val, const float *x,
                 float *y, float alpha, float beta, int lanes, void *stream);

/* y = a*x + b*y (VecAXPBY in linear/deflated_fgmres.py:174-295). */
int ssx_axpby_f64(int64_t n, double a, const double *x, double b, double *y, void *stream);
int ssx_axpby_f32(int64_t n, float a, const float *x, float b, float *y, void *stream);
/* Precision casts at the FP32 preconditioner boundary. */
int ssx_cast_f64_f32(int64_t n, const double *x, float *y, void *stream);
int ssx_cast_f32_f64(int64_t n, const float *x, double *y, void *stream);

/* out[j] = V[:, j] . x for j < k (column stride ld, k <= 65535): the classical Gram-Schmidt
 * projections of one Arnoldi step (linear/deflated_fgmres.py:261-263, cython/_kernels.pyx:190-192).
 * part: scratch of at least 64 * k doubles (per-chunk partials, summed in a fixed order). */
int ssx_dot_cols(int64_t n, int k, const double *V, int64_t ld, const double *x, double *part, double *out,
                 void *stream);
/* out[0] = x . y, deterministic two-stage reduction; part: >= 1024 doubles of scratch
 * (cython/_kernels.pyx:100-123 `dot`/`norm2`). */
int ssx_dot(int64_t n, const double *x, const double *y, double *part, double *out, void *stream);
/* x += sign * V[:, :k] h (the CGS update and the Krylov solution update). */
int ssx_gemv_cols(int64_t n, int k, const double *V, int64_t ld, const double *h, double sign, double *x,
                  void *stream);

/* Y_j = A X_j for k vectors stored as rows (X_j = X + j*ldx, Y_j = Y + j*ldy); the matrix is read
 * once per 8 vectors.  Replaces the per-column PETSc MatMult of linear/orthogonalize.py:139-140. */
int ssx_spmm_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, int k,
                 const double *X, int64_t ldx, double *Y, int64_t ldy, int lanes, void *stream);
/* A-orthonormalisation of the deflation basis B (kmax rows of length n, row stride ldb, oldest
 * first), newest column first with the drop rule |v . A b| <= tol (linear/orthogonalize.py:124-170):
 * W[i], AW[i] = w and A w of column kmax-1-i, sgn[i] = sign of its A-norm (0 = dropped, W[i] = 0).
 * work: scratch of at least 3*kmax + max(64*kmax, 1024) doubles.  No host round trip. */
int ssx_a_orthogonalize(int64_t n, int kmax, const double *B, int64_t ldb, const int32_t *indptr,
                        const int32_t *indices, const double *val, int lanes, double tol, double *W, double *AW,
                        double *sgn, double *work, void *stream);

/* The same A-orthonormalisation with no scratch argument and no allocation: the scratch is a
 * file-scope device array of the library (34,304 bytes, one copy per GPU, kept until the process
 * ends), so kmax is limited to 64 (larger kmax is rejected with status 1).  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  Calls that overlap on one GPU race on the shared scratch; the scratch needs
 * no clearing between calls (every word is written before it is read). */
int ssx_a_orthogonalize_noalloc(int64_t n, int kmax, const double *B, int64_t ldb, const int32_t *indptr,
                                const int32_t *indices, const double *val, int lanes, double tol, double *W,
                                double *AW, double *sgn, void *stream);

/* One fused Chebyshev(Jacobi) smoothing step on CSR A: r = b - A x_in (x_in NULL: zero), d = a d +
 * c dinv r (first != 0: d = c dinv r), x_out = x_in + d; x_out must differ from x_in.  Replaces the
 * residual, smoother and update calls of one step of linear/solver.py:155-162,417-436. */
int ssx_cheb_step_f32(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const float *val,
                      const float *b, const float *x_in, float *x_out, float *d, const float *dinv, float c, float a,
                      int first, int lanes, void *stream);
int ssx_cheb_step_f64(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val,
                      const double *b, const double *x_in, double *x_out, double *d, const double *dinv, double c,
                      double a, int first, int lanes, void *stream);
/* Chebyshev/Jacobi smoother step: d = a*d + c*dinv.*r; x += d (replaces the Richardson/SOR and
 * Chebyshev/Jacobi smoother applies of linear/solver.py:155-162,417-436). */
int ssx_cheb_update_f32(int64_t n, const float *dinv, const float *r, float c, float a, float *x, float *d,
                        void *stream);
int ssx_cheb_update_f64(int64_t n, const double *dinv, const double *r, double c, double a, double *x, double *d,
                        void *stream);
/* dinv[r] = 1 / A[r, r + col0] (0 where the diagonal is zero). */
int ssx_inv_diag_f32(int64_t n_rows, int64_t col0, const int32_t *indptr, const int32_t *indices, const float *val,
                     float *dinv, void *stream);
int ssx_inv_diag_f64(int64_t n_rows, int64_t col0, const int32_t *indptr, const int32_t *indices, const double *val,
                     double *dinv, void *stream);

/* P4 tetrahedra (35 nodes, 24 points): element tangent matrices ke (n_elem x 105 x 105, row-major,
 * row = local dof 3*node+comp) = sum_q B^T (w D) B on the FP64 matrix cores (v_mfma_f64_16x16x4).
 * With shear/lame (per point, both or neither) D is the regularised (1-r) DS + r D_el, D_el the
 * elastic tangent (lame vol vol^T + 2 G ident), which folds K_r = (1-r) K + r K_el into the element
 * matrices (constitutive/problem.py:1595-1644).
 * Replaces the element stage of assemble_tangent_values_3d_p2_c (cython/assemble_tangent_values_3d.c:45-94). */
int ssx_element_matrices_p4(const double *dphi, const double *DS, const double *w, int64_t n_elem, double r,
                            const double *shear, const double *lame, double *ke, void *stream);

/* P1/P2 tetrahedra (n_p 4 or 10): element tangent matrices ke (n_elem x 3n_p x 3n_p, row-major):
 * persistent wavefronts with the next element's point data in flight for the 1-point P1 rule, one
 * wavefront per element otherwise (P2 elements at any rule, e.g. the P2 level of the P4 multigrid
 * at the 24 P4 points); optional K_r fold exactly as ssx_element_matrices_p4.  Replaces the
 * element stage of cython/assemble_tangent_values_3d.c:13-109. */
int ssx_element_matrices_3d(int n_p, const double *dphi, const double *DS, const double *w, int64_t n_elem, int n_q,
                            double r, const double *shear, const double *lame, double *ke, void *stream);

/* Node gather of element matrices into the fixed CSR pattern: one wavefront per mesh node sums the
 * node's dim rows over its elements (node slots ns_ptr/ns_elem/ns_lnode, element order) at the
 * int16 column offsets ns_off (n_slots x nld, relative to the row start; a constrained column
 * holds the row length L),
 * sets constrained rows to their unit diagonal, optionally blends (1-r) K + r elastic, and writes
 * FP64 values and an optional FP32 copy.  nodes = the node ids to gather (n_nodes of them; NULL:
 * nodes 0 .. n_nodes-1), max_len = longest free row among them (< 32768; it sizes the LDS of a
 * workgroup, so callers group nodes of similar row length).  dim 3, nld 12/30/105.
 * Together with ssx_element_matrices_3d / _p4 replaces assemble_tangent_values_3d_rows_c
 * (cython/assemble_tangent_values_3d.c:111-214) and the constrained-diagonal step of
 * assemble_owned_tangent_values (fem/distributed_tangent.py:1559-1612). */
int ssx_gather_node_rows(int dim, int nld, const double *ke, const int32_t *ns_ptr, const int32_t *ns_elem,
                         const uint8_t *ns_lnode, const int16_t *ns_off, const int32_t *indptr,
                         const uint8_t *row_free, const int32_t *nodes, int64_t n_nodes, int max_len, double r,
                         const double *elastic, double *out, float *out32, void *stream);

/* P2 tetrahedra (n_p 10, 11-point rule): the upper node-pair blocks of the element tangent
 * matrices, kb (n_elem x 55 x 3 x 3): pair p = (a, b), a <= b, in row-major order over the upper
 * triangle (p = a*10 - a(a-1)/2 + b - a), kb[e][p][i][j] = K_e[3a+i, 3b+j] = sum_q w B_a^T D B_b.
 * The blocks below the diagonal are the transposes (the return mapping's tangent is symmetric).
 * Optional K_r fold exactly as ssx_element_matrices_p4 (shear/lame given).  Optional branch ids per
 * point (int8, as written by ssx_constitutive_3d for this DS; needs shear/lame): a point with id 0
 * (elastic) has DS = D_el(shear, lame), and its contribution is formed in closed form.  Replaces
 * the element stage of cython/assemble_tangent_values_3d.c:13-109. */
int ssx_element_blocks_3d(int n_p, const double *dphi, const double *DS, const double *w, int64_t n_elem, int n_q,
                          double r, const double *shear, const double *lame, const int8_t *branch, double *kb,
                          void *stream);

/* Node gather of the node-pair blocks of ssx_element_blocks_3d into the fixed CSR pattern, with
 * the node slots, offsets, constrained-row, blend and FP32-copy semantics of ssx_gather_node_rows
 * (n_p 10; max_len <= 2729, one LDS copy of a node's three rows per workgroup).  nodes = the
 * order in which the nodes are gathered (NULL: 0 .. n_nodes-1); the result does not depend on it.
 * Together with ssx_element_blocks_3d replaces assemble_tangent_values_3d_rows_c
 * (cython/assemble_tangent_values_3d.c:111-214) and the constrained-diagonal step of
 * assemble_owned_tangent_values (fem/distributed_tangent.py:1559-1612). */
int ssx_gather_node_blocks(int n_p, const double *kb, const int32_t *ns_ptr, const int32_t *ns_elem,
                           const uint8_t *ns_lnode, const int16_t *ns_off, const int32_t *indptr,
                           const uint8_t *row_free, const int32_t *nodes, int64_t n_nodes, int max_len, double r,
                           const double *elastic, double *out, float *out32, void *stream);

/* ---- P2 tetrahedra with a constant Jacobian (straight edges, edge nodes at the midpoints) ----
 * The three entry points below are ssx_strain, ssx_force_elem and ssx_element_blocks_3d for such
 * meshes, with the per-point dphi and w replaced by per-element factors:
 *   gm  (n_elem, 10): (J^-1)[d][k] at 3 d + k (the cofactor inverse of fem/assembly.py:139-245,
 *                     taken at the rule's first point) and |det J| at 9;
 *   ref (3 n_q n_p + n_q): the rule's reference gradients h_k(q, n) at (k n_q + q) n_p + n, then
 *                     its weights wf[q];
 * so that dphi_d(e, q, n) = sum_k gm[e][3 d + k] h_k(q, n) and w(e, q) = gm[e][9] wf[q].  Results
 * agree with the dphi forms to rounding.  No counterpart in the reference, whose kernels take dphi
 * and w (cython/assemble_tangent_values_3d.c:13-109, :216-273, :275-327); for callers that own the
 * mesh and hold it as straight-sided P2 tetrahedra.  n_p 10, n_q 11 only. */
int ssx_strain_affine_3d(int n_p, const double *gm, const double *ref, const int32_t *edofs, const double *u,
                         int64_t n_elem, int n_q, double *E, void *stream);
int ssx_force_elem_affine_3d(int n_p, const double *gm, const double *ref, const double *S, int64_t n_elem, int n_q,
                             double *fe, void *stream);
/* ds_sym = 1: DS is the packed (n_int x 21) tangent of ssx_constitutive_3d_sym. */
int ssx_element_blocks_affine_3d(int n_p, const double *gm, const double *ref, const double *DS, int ds_sym,
                                 int64_t n_elem, int n_q, double r, const double *shear, const double *lame,
                                 const int8_t *branch, double *kb, void *stream);

/* ---- learned continuation predictor (no reference counterpart: it fills the predictor slot of
 * continuation/indirect.py:2703-2830, where the reference uses the secant of :550-566) ---- */

/* C (M x N fp32, row stride ldc) = op(A) op(B) [+ bias[j]] [tanh] [+ beta*C]; bf16 operands given
 * by element strides (A[i*a_si + k*a_sk], B[k*b_sk + j*b_sj]); optional bf16 copy Cb (same ldc);
 * act 0 none, 1 tanh.  bf16 MFMA with FP32 accumulation. */
int ssx_gemm_bf16(int64_t M, int64_t N, int64_t K, const void *A, int64_t a_si, int64_t a_sk, const void *B,
                  int64_t b_sk, int64_t b_sj, float *C, int64_t ldc, const float *bias, int act, void *Cb, float beta,
                  void *stream);
/* dZ = dY (1 - Y^2) with optional bf16 copy. */
int ssx_tanh_bwd(int64_t n, const float *dY, const float *Y, float *dZ, void *dZb, void *stream);
/* db[j] = sum_i dZ[i*N + j]. */
int ssx_colsum_f32(int64_t M, int64_t N, const float *dZ, float *db, void *stream);
/* Adam step (step >= 1) on FP32 weights w with moments m, v; optional bf16 copy wb. */
int ssx_adam_f32(int64_t n, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                 int step, void *wb, void *stream);
int ssx_f32_to_bf16(int64_t n, const float *x, void *y, void *stream);
int ssx_bf16_to_f32(int64_t n, const void *x, float *y, void *stream);
/* Masked MSE: g = 2 (p - t) mask * inv_count; loss partial sums per 256 entries in loss_part. */
int ssx_mse_f32(int64_t n, const float *p, const float *t, const float *mask, float inv_count, float *g,
                float *loss_part, void *stream);

/* Dense Galerkin coarse operator Ac (n_c x n_c, column-major, caller-zeroed) = P^T A P with P given
 * as CSR rows (p_ptr, p_idx, p_w) over the fine rows of A (nnz entries).  The products are summed in
 * 64-bit fixed point with integer atomics, so the result is bitwise the same at every call; the
 * scale is 2^(61 - ceil(log2(max|a| * w_bound))), where w_bound must bound sum_ij |p_iI| |p_jJ| over
 * the products of any one coarse entry (e.g. the largest column sum of |P| times the longest row of
 * A times max|p|).  work: 8 bytes of device scratch.  Setup-time work.
 * Replaces the PETSc MatPtAP of the p-multigrid set-up (linear/solver.py:510-513). */
int ssx_galerkin_dense(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, int64_t nnz,
                       const int32_t *p_ptr, const int32_t *p_idx, const double *p_w, int64_t n_c, double w_bound,
                       double *Ac, void *work, void *stream);

/* Sparse Galerkin coarse operator onto a given coarse CSR pattern (c_indptr, c_indices sorted per
 * row, n_c rows, c_nnz entries, rows of at most max_c_row <= 1024 entries): c_val = P^T A P at the
 * positions the pattern holds, products outside it dropped (for nested P1-in-P2 spaces they sum to
 * zero).  pt_ptr/pt_idx/pt_w is P^T in CSR (n_c rows): each coarse row is gathered by one
 * wavefront, every entry written once.  Fixed-point sums as in ssx_galerkin_dense (bitwise
 * repeatable, the same bits as the scatter form).  Replaces the same PETSc MatPtAP
 * (linear/solver.py:510-513) for coarse levels too large for a dense matrix. */
int ssx_galerkin_csr(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val, int64_t nnz,
                     const int32_t *p_ptr, const int32_t *p_idx, const double *p_w, const int32_t *pt_ptr,
                     const int32_t *pt_idx, const double *pt_w, int64_t n_c, const int32_t *c_indptr,
                     const int32_t *c_indices, int64_t c_nnz, int max_c_row, double w_bound, double *c_val,
                     void *work, void *stream);

/* Coarse solve y = Ainv x of the p-multigrid with the explicit inverse of the symmetric coarse
 * operator kept as the upper triangle of 64 x 64 tiles (half the bytes of a dense product per
 * apply; deterministic).  Replaces the coarse solve of the manual V-cycle (linear/solver.py:97-160,
 * Hypre there).  Sizes: tiles = ceil(n/64)(ceil(n/64)+1)/2 * 4096 elements, part = ceil(n/64)^2 * 64. */
int ssx_sym_tiles_sizes(int64_t n, int64_t *tile_elems, int64_t *part_elems);
/* Pack a symmetric n x n matrix (FP64 if a_f64, else FP32) into tiles (FP64 if t_f64, else FP32). */
int ssx_sym_tiles_pack(int64_t n, const void *A, int a_f64, void *tiles, int t_f64, void *stream);
/* Explicit inverse of a symmetric positive definite n x n matrix A (FP64, symmetric, either storage
 * order) into the symmetric tile store above (tiles in FP64 if t_f64, else FP32), computed in FP64
 * (f64 = 1) or FP32: blocked Cholesky A = L L^T on 64 x 64 tiles, W = L^-1, A^-1 = W^T W for the
 * stored upper tiles; the matrix is padded to a multiple of 64 with the identity, and the tiles are
 * zero outside n x n.  work: ssx_spd_inverse_sizes elements of the computation precision.  *info
 * (device int) receives 0, or 1 + the index of the first non-positive pivot.  Replaces the dense
 * coarse factorisation of the p-multigrid set-up (Hypre/PETSc there, linear/solver.py:466-513). */
int ssx_spd_inverse_sizes(int64_t n, int64_t *work_elems);
int ssx_spd_inverse_tiles(int64_t n, const double *A, int f64, void *work, void *tiles, int t_f64, int *info,
                          void *stream);
int ssx_sym_tiles_gemv_f32(int64_t n, const float *tiles, const float *x, float *part, float *y, void *stream);
int ssx_sym_tiles_gemv_f64(int64_t n, const double *tiles, const double *x, double *part, double *y, void *stream);

/* ---- device probe (no reference counterpart: measurement only) ---- */

/* Runs `blocks` workgroups of 256 threads, each lane `iters` rounds of independent FP64 dependency
 * chains: mode 0 vector FMA, mode 1 v_mfma_f64_16x16x4 (four chains per wavefront).  out receives
 * blocks*256 doubles (keeps the work alive); *flops (host) the FP64 operations launched, so
 * flops / elapsed time is the rate this GPU attains. */
int ssx_probe_fp64(int mode, int64_t blocks, int64_t iters, double *out, double *flops, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* SSX_H */
