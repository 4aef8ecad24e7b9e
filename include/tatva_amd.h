/* tatva_amd: C ABI of the MI355X (gfx950) element-assembly / colored-Hessian /
 * matrix-free-HVP / Jacobi-PCG path.
 *
 * Conventions
 *   - Every function returns an int status (TATVA_OK == 0); on failure the
 *     message is available from tatva_last_error() (thread-local).
 *   - Device buffers are plain pointers owned by the caller, sizes are element
 *     counts.  Compute calls are asynchronous on the caller's hipStream_t
 *     (passed as void*), except the PCG solvers, which return after the solve.
 *   - Vectors are node-major interleaved: dof = node * ncomp + component, as in
 *     tatva's flat state vectors.
 *   - "dof_map" arguments map a full dof to its reduced (free) index, or -1 for a
 *     constrained dof (Lifter/Fixed); NULL means the identity.
 *   - CSR arrays are int32 (nnz < 2^31), sorted column indices per row.
 *   - One mesh handle must not be used concurrently from two streams: it owns
 *     the element scratch buffer and reduction counters its kernels share.
 *
 * Each declaration cites the reference function (tatva v0.11.1, path relative
 * to the tatva repository) whose behaviour it replaces.
 */
#ifndef TATVA_AMD_H
#define TATVA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  TATVA_OK = 0,
  TATVA_EINVAL = 1,   /* invalid argument (mirrors ValueError / TypeError)   */
  TATVA_EHIP = 2,     /* HIP runtime error                                    */
  TATVA_ENOMEM = 3,   /* allocation failed / workspace too small              */
  TATVA_EBREAKDOWN = 4, /* CG breakdown: p.Ap <= 0 or non-finite              */
  TATVA_ENOTCONV = 5  /* iteration limit reached before the tolerance         */
};

enum { TATVA_TRI3 = 0, TATVA_TET4 = 1, TATVA_HEX8 = 2 };

enum {
  TATVA_DENSITY_POISSON = 1,        /* psi = |grad u|^2/2 - f u        (p0 = f)           */
  TATVA_DENSITY_LINEAR_ELASTIC = 2, /* psi = sigma:eps/2               (p0 = mu, p1 = lambda) */
  TATVA_DENSITY_NEO_HOOKEAN = 3     /* psi = mu/2(I1-d-2lnJ)+lambda/2 ln^2 J (p0 = mu, p1 = lambda) */
};

/* Built-in energy density descriptor (replaces the user JAX density callables of
 * tests/test_sparse.py:20-38 and tests/test_sparse_tracer.py:103-115).
 * elem_params != NULL gives every element its own parameters (device array,
 * n_elements x np in element order, np = 1 for Poisson (f), 2 otherwise (mu,
 * lambda)); p0.. are then unused.  This is how element quantities reach a
 * density (Operator.map(..., element_quantity=...), tatva/operator.py:229,256).
 * points_per_element: tatva_density_eval with elem_params only (its points are
 * consecutive per element). */
typedef struct {
  int32_t kind;
  int32_t points_per_element;
  double p0, p1, p2;
  const double *elem_params;
} tatva_density;

typedef struct tatva_mesh tatva_mesh;

/* PCG result. status: TATVA_OK, TATVA_EBREAKDOWN or TATVA_ENOTCONV. */
typedef struct {
  int32_t iterations;
  int32_t status;
  double rel_residual;    /* ||r_k|| / ||b|| of the recursively updated residual */
  double b_norm;
  double op_ms;           /* with profiling on: summed duration of the sampled operator launches */
  int32_t op_samples;     /* number of sampled operator launches (0 with profiling off) */
  int32_t reserved;
  double op_main_ms;      /* with profiling on: the sampled operators' main kernel alone (the
                             SpMV; the tile kernel of a matrix-free HVP, without its gather) */
} tatva_pcg_stats;

const char *tatva_last_error(void);
int tatva_version(void);
int tatva_device_count(int *out);

/* Profiling switch (no counterpart in the reference): when on, each captured
 * PCG chunk brackets its first operator application (SpMV or matrix-free HVP)
 * with HIP events on the solver stream and reports the summed duration in
 * tatva_pcg_stats.op_ms / op_samples. */
int tatva_set_profiling(int enable);

/* Process-wide options (no counterpart in the reference).
 * TATVA_OPT_MATFREE_LIN_CACHE: whether tatva_pcg_matfree_f32 linearises the
 * density once per solve at its fixed state (Neo-Hookean: F^-1 and
 * mu - lambda ln J per quadrature point, FP32, in the workspace) so that every
 * HVP of the solve reads it instead of recomputing it (same arithmetic, same
 * result): 0 (default) never, 1 for Hex8 meshes, 2 always.
 * TATVA_OPT_MATFREE_GEOM_CACHE: whether tatva_pcg_matfree_f32 on an affine
 * Hex8 mesh (every element a parallelepiped) writes each element's Jacobian
 * inverse (FP32) and det J (FP64) once per solve, 44 bytes per element in the
 * workspace, so that every HVP reads them instead of the node coordinates (the
 * values the product computes itself, so the same result bit for bit): 1
 * (default) on affine Hex8 meshes, 0 never.  Both caches are in the workspace
 * only when selected: set the options before querying its size.
 * TATVA_OPT_MATFREE_DEFERRED_UPDATE: whether tatva_pcg_matfree_f32 leaves each
 * CG iteration's p and x updates to the tile kernel of the next HVP (which has
 * memory bandwidth to spare) instead of its step kernel, with z in two
 * alternating buffers: 1 (default) on, 0 off.  The iterates are the same bits
 * either way.
 * TATVA_OPT_MATFREE_FULL_VECTORS: whether tatva_pcg_matfree_f32 with a dof map
 * runs its CG on full-length vectors (every dof of the mesh, the inverse
 * diagonal 0 outside the system), so that no kernel of the iteration reads the
 * map: b, x and the diagonal go in and x comes out through it once per solve.
 * 1 (default) on, 0 off.  The products are the same bits; the dot products sum
 * in another order.  The workspace then holds full-length vectors: set the
 * option before querying its size. */
enum {
  TATVA_OPT_MATFREE_LIN_CACHE = 1,
  TATVA_OPT_MATFREE_GEOM_CACHE = 2,
  TATVA_OPT_MATFREE_DEFERRED_UPDATE = 3,
  TATVA_OPT_MATFREE_FULL_VECTORS = 4
};
int tatva_set_option(int32_t option, int32_t value);

/* ---------------------------------------------------------------- mesh ---- */

/* Uploads a mesh (host arrays) and builds its node->element incidence.
 * Replaces Mesh(coords, elements) (tatva/mesh.py:53-85) + Operator.__post_init__
 * / __check_init__ validation (tatva/operator.py:112-170).  quad_points (nq*dim)
 * and quad_weights (nq) give the element quadrature (element/base.py:37-51). */
int tatva_mesh_create(int device, int elem_kind, const double *coords, int64_t n_nodes,
                      int dim, const int32_t *elements, int64_t n_elements, int nne, int nq,
                      const double *quad_points, const double *quad_weights,
                      tatva_mesh **out);
int tatva_mesh_destroy(tatva_mesh *mesh);

/* Sizes of a mesh handle's device layout (no counterpart in the reference; for
 * byte counts of the kernels): out[0..n) of {n_nodes, n_elements, n_tiles,
 * n_slots (tile, local node) pairs, largest local node count of a tile, largest
 * element color count of a tile, affine Hex8 flag, quadrature points}. */
int tatva_mesh_info(const tatva_mesh *mesh, int64_t *out, int32_t n);

/* Integration weights detJ*w_q, out (E, Q).
 * Replaces Operator.get_integration_weights (tatva/operator.py:172-192). */
int tatva_weights(tatva_mesh *mesh, double *out, void *stream);

/* Values at quadrature points, out (E, Q, ncomp) from nodal u (N, ncomp).
 * Replaces Operator.eval (tatva/operator.py:358-377). */
int tatva_eval(tatva_mesh *mesh, int ncomp, const double *u, double *out, void *stream);

/* Gradients at quadrature points, out (E, Q, ncomp, dim), out[...,i,j] = du_i/dx_j.
 * Replaces Operator.grad (tatva/operator.py:379-397; element/base.py:99-115). */
int tatva_grad(tatva_mesh *mesh, int ncomp, const double *u, double *out, void *stream);

/* Integral of values given at quadrature points (E, Q, ncomp), or of a nodal
 * field (N, ncomp) when nodal != 0.  per_element (E, ncomp) and total (ncomp)
 * may be NULL.  Replaces Operator.integrate / integrate_per_element /
 * _integrate_quad_array (tatva/operator.py:307-356). */
int tatva_integrate(tatva_mesh *mesh, int ncomp, int nodal, const double *values,
                    double *per_element, double *total, void *stream);

/* Density values psi at n_points quadrature points from gradients
 * (n_points, ncomp, dim) and, for Poisson, values (n_points).  Replaces the
 * autovmap'ed density callables evaluated on Operator.grad output
 * (tests/test_sparse.py:33-38). */
int tatva_density_eval(int dim, int64_t n_points, const tatva_density *density,
                       const double *grad, const double *values, double *psi, void *stream);

/* Total energy E(u) = sum_e sum_q psi(grad u) detJ w - f_ext.u (f_ext may be NULL),
 * written to out[0].  Replaces op.integrate(psi(op.grad(u))) (operator.py:307-319). */
int tatva_energy(tatva_mesh *mesh, const tatva_density *density, const double *u,
                 const double *f_ext, double *out, void *stream);

/* Residual r = dE/du (jax.grad of the energy; the reference's scatter-add is
 * the transpose of the gather at tatva/operator.py:221).  Deterministic node
 * gather, no float atomics.  With dof_map, writes only free entries to
 * r[dof_map[dof]] (Lifter.reduce_adjoint for Fixed, lifter/base.py:235-251). */
int tatva_residual(tatva_mesh *mesh, const tatva_density *density, const double *u,
                   const double *f_ext, const int32_t *dof_map, double *r, void *stream);

/* Hessian-vector product y = H(u) v, FP64 (jax.jvp(jax.grad(E), (u,), (v,))[1],
 * README.md:93; sparse/base.py:213,264).  u is a full vector; v and y are
 * reduced vectors when dof_map != NULL (v lifted with zeros on constrained dofs). */
int tatva_hvp(tatva_mesh *mesh, const tatva_density *density, const double *u,
              const double *v, const int32_t *dof_map, double *y, void *stream);

/* Same in FP32 (FP32 state, tangent and accumulation; geometry from FP64 coords). */
int tatva_hvp_f32(tatva_mesh *mesh, const tatva_density *density, const float *u,
                  const float *v, const int32_t *dof_map, float *y, void *stream);

/* Diagonal of H(u) (Jacobi preconditioner), FP64 / FP32 output (reduced with dof_map). */
int tatva_hessian_diag(tatva_mesh *mesh, const tatva_density *density, const double *u,
                       const int32_t *dof_map, double *diag, void *stream);
int tatva_hessian_diag_f32(tatva_mesh *mesh, const tatva_density *density, const float *u,
                           const int32_t *dof_map, float *diag, void *stream);

/* ------------------------------------------------------ colored Hessian ---- */

/* Workspace bytes for tatva_colored_hessian on this mesh with ncomp components. */
int64_t tatva_colored_hessian_workspace(tatva_mesh *mesh, int ncomp);

/* Colored sparse Hessian decompressed into CSR data.  Replaces sparse.jacfwd /
 * colored_jacobian_batch (tatva/sparse/base.py:139-176,230-270): one JVP per
 * color with seed (colors == c), J_compressed[row, color], then
 * data[k] = J_compressed[rows[k], colors[indices[k]]] (sparse/base.py:167-174).
 * The pattern (indptr, indices) and colors refer to the reduced (free) dofs
 * when dof_map != NULL; red_to_full (n_rows) then lists the full dof of each row. */
int tatva_colored_hessian(tatva_mesh *mesh, const tatva_density *density, const double *u,
                          const int32_t *dof_map, const int32_t *red_to_full, int64_t n_rows,
                          const int32_t *colors, int32_t n_colors, const int32_t *indptr,
                          const int32_t *indices, int64_t nnz, double *data, void *workspace,
                          int64_t workspace_bytes, void *stream);

/* The same for a general lifter (Fixed and Periodic, lifter/base.py:201-251,
 * constraints.py:184-220): the reduced Hessian P^T H P with P the linear part of
 * the lift.  read_map (full dofs) gives each full dof's reduced column (or -1);
 * reduced row r sums the full rows row_list[row_ptr[r] .. row_ptr[r+1]).  With
 * flags & 1 every element's columns are accumulated one at a time (needed when an
 * element holds two dofs of one reduced column, e.g. a slave and its master). */
int tatva_colored_hessian_mapped(tatva_mesh *mesh, const tatva_density *density,
                                 const double *u, const int32_t *read_map,
                                 const int32_t *row_ptr, const int32_t *row_list, int64_t n_rows,
                                 const int32_t *colors, int32_t n_colors, const int32_t *indptr,
                                 const int32_t *indices, int64_t nnz, int32_t flags, double *data,
                                 void *workspace, int64_t workspace_bytes, void *stream);

/* L2 projection (Operator.project, tatva/operator.py:523-554 and the mass /
 * right-hand-side forms of tatva/utils.py:232-262).  tatva_mass_matrix writes
 * the consistent mass matrix of an ncomp-component field (block diagonal over
 * components, sum_q N_a N_b detJ w) into CSR data on a (reduced) pattern, as
 * tatva_colored_hessian does; tatva_project_rhs writes
 * out[node, c] = sum_e sum_q N_a(xi_q) vals[e, q, c] detJ w_q (full nodes). */
int tatva_mass_matrix(tatva_mesh *mesh, int32_t ncomp, const int32_t *dof_map,
                      const int32_t *red_to_full, int64_t n_rows, const int32_t *colors,
                      int32_t n_colors, const int32_t *indptr, const int32_t *indices,
                      int64_t nnz, double *data, void *workspace, int64_t workspace_bytes,
                      void *stream);
int tatva_project_rhs(tatva_mesh *mesh, int32_t nc, const double *quad_values, double *out,
                      void *stream);
/* The projection's mass matrix under a general lifter (Operator.project with a
 * Periodic constraint, tatva/utils.py:225-246): the Jacobian of the lifted
 * bilinear form, rows of the free test dofs, columns through the lift's read map
 * (a slave's column added to its master's): E_free^T M P.  Reduced row r sums
 * the full rows row_list[row_ptr[r] .. row_ptr[r+1]) (the projection passes its
 * own free dof only); flags & 1 as in tatva_colored_hessian_mapped.  Not
 * symmetric in general: solve it with tatva_bicgstab_csr. */
int tatva_mass_matrix_mapped(tatva_mesh *mesh, int32_t ncomp, const int32_t *read_map,
                             const int32_t *row_ptr, const int32_t *row_list, int64_t n_rows,
                             const int32_t *colors, int32_t n_colors, const int32_t *indptr,
                             const int32_t *indices, int64_t nnz, int32_t flags, double *data,
                             void *workspace, int64_t workspace_bytes, void *stream);

/* Operator.interpolate (tatva/operator.py:399-468 with find_containing_polygons,
 * tatva/mesh.py:295-390) on a 2-D Tri3 mesh: out[p, c] = nodal values (N x nc)
 * interpolated at points (n_points x 2); the containing element is the first
 * whose vertex polygon holds the point (edges included), mapped to reference
 * coordinates by one Newton step from the first quadrature point.  Points in no
 * element get NaN; their number is returned in *n_outside (synchronises stream).
 * The first call on a mesh builds its point-location grid (elements binned by
 * bounding box, kept with the mesh until tatva_mesh_destroy); each point then
 * tests only the elements of its grid cell, in ascending order, which finds
 * the same element as the reference's scan over all of them. */
int tatva_interpolate(tatva_mesh *mesh, int64_t n_points, const double *points, int32_t nc,
                      const double *nodal, double *out, int32_t *n_outside, void *stream);

/* rows[k] = row of entry k and col_colors[k] = colors[indices[k]] for a CSR
 * pattern (compute_rows_cols, tatva/sparse/base.py:108-136). */
int tatva_rows_cols(int64_t n, const int32_t *indptr, const int32_t *indices,
                    const int32_t *colors, int32_t *rows, int32_t *col_colors, void *stream);

/* ------------------------------------------------------- CSR and solvers -- */

/* y = A x for a CSR matrix (FP64).  Replaces ColoredMatrix.to_bcsr() @ x
 * (tatva/sparse/base.py:95-100). */
int tatva_spmv_csr(int64_t n, int64_t nnz, const int32_t *indptr, const int32_t *indices,
                   const double *data, const double *x, double *y, void *stream);

/* Node-block view of a CSR pattern (no counterpart in the reference): when the
 * rows come in groups of block_size (one node) sharing one column list of
 * aligned block_size-wide blocks, writes one block column per block into
 * block_cols (nnz / block_size^2 entries) and sets *ok = 1; otherwise *ok = 0.
 * The CSR arrays are not changed.  Setup call: returns after the check. */
int tatva_csr_block_view(int64_t n, int64_t nnz, const int32_t *indptr, const int32_t *indices,
                         int32_t block_size, int32_t *block_cols, int32_t *ok, void *stream);

/* y = A x through the node-block view (same result as tatva_spmv_csr). */
int tatva_spmv_bcsr(int64_t n, const int32_t *indptr, const int32_t *block_cols,
                    int32_t block_size, const double *data, const double *x, double *y,
                    void *stream);

/* Jacobi-PCG on a CSR matrix in FP64 (no counterpart in the reference; it
 * replaces the user-side linear solve, cf. spsolve in tatva/utils.py:186-198).
 * x holds x0 on entry.  Iterations run in hipGraph-captured chunks of `chunk`.
 * With block_cols (from tatva_csr_block_view) the operator reads the CSR data
 * through the node-block view; without, the call first copies the matrix into
 * a sliced-ELL layout (64-row slices) in the caller's workspace (one host
 * synchronisation per call).  The workspace query takes the HOST copy of
 * indptr when the solve will run without a block view (NULL: with one), and
 * then includes the sliced-ELL copy.  The library keeps no buffer of the
 * solve after it returns.  The PCG workspaces (this one and the matrix-free
 * ones below) must be 256-byte aligned (device allocations are): the CG step
 * moves its vectors in 16-byte loads from 256-byte offsets of it. */
int64_t tatva_pcg_csr_workspace(int64_t n, const int32_t *host_indptr);
int tatva_pcg_csr(int64_t n, int64_t nnz, const int32_t *indptr, const int32_t *indices,
                  const double *data, const int32_t *block_cols, int32_t block_size,
                  const double *b, double *x, double tol, int32_t maxiter, int32_t chunk,
                  void *workspace, int64_t workspace_bytes, tatva_pcg_stats *stats,
                  void *stream);

/* Jacobi-preconditioned BiCGSTAB on a general CSR matrix in FP64, x in place (x0
 * on entry), until ||b - A x|| <= tol ||b||; stats as tatva_pcg_csr.  It replaces
 * the reference's spsolve of a non-symmetric projection matrix
 * (tatva/utils.py:186-198).  Host-driven: returns after the solve. */
int64_t tatva_bicgstab_csr_workspace(int64_t n);
int tatva_bicgstab_csr(int64_t n, int64_t nnz, const int32_t *indptr, const int32_t *indices,
                       const double *data, const double *b, double *x, double tol,
                       int32_t maxiter, void *workspace, int64_t workspace_bytes,
                       tatva_pcg_stats *stats, void *stream);

/* Jacobi-PCG with the matrix-free FP32 Hessian-vector product of (mesh, density)
 * at the FP32 state u (full vector); b, x, diag are reduced FP32 vectors (n
 * entries, dof_map: full dof -> reduced index, -1 on a constrained dof, each
 * reduced index read by one full dof).  The workspace holds the CG vectors
 * (full-length under TATVA_OPT_MATFREE_FULL_VECTORS, the default), a copy of u
 * and the caches the options select: for a Neo-Hookean density its
 * linearisation at u (TATVA_OPT_MATFREE_LIN_CACHE), for an affine Hex8 mesh its
 * element geometry (TATVA_OPT_MATFREE_GEOM_CACHE). */
int64_t tatva_pcg_matfree_workspace(tatva_mesh *mesh, const tatva_density *density, int64_t n);
int tatva_pcg_matfree_f32(tatva_mesh *mesh, const tatva_density *density, const float *u,
                          const int32_t *dof_map, int64_t n, const float *diag, const float *b,
                          float *x, double tol, int32_t maxiter, int32_t chunk, void *workspace,
                          int64_t workspace_bytes, tatva_pcg_stats *stats, void *stream);

/* The same solve when constraints tie several full dofs to one reduced dof
 * (Periodic, lifter/constraints.py:184-220): the operator is P^T H(u) P with P
 * the linear part of the lift (Lifter.lift_from_zeros, lifter/base.py:201-229)
 * and P^T its adjoint (Lifter.reduce_adjoint, lifter/base.py:235-251), as the
 * reference's jax.jvp of the reduced residual computes it.  read_map gives each
 * full dof the reduced index it reads (-1: constraint value); rows_ptr/rows_list
 * (n + 1, CSR) list the full dofs summed into each reduced index, ascending. */
int64_t tatva_pcg_matfree_mapped_workspace(tatva_mesh *mesh, int64_t n);
int tatva_pcg_matfree_f32_mapped(tatva_mesh *mesh, const tatva_density *density, const float *u,
                                 const int32_t *read_map, const int32_t *rows_ptr,
                                 const int32_t *rows_list, int64_t n, const float *diag,
                                 const float *b, float *x, double tol, int32_t maxiter,
                                 int32_t chunk, void *workspace, int64_t workspace_bytes,
                                 tatva_pcg_stats *stats, void *stream);

/* ------------------------------------------------------- batched PCG ------ */

/* A batch runs many matrix-free FP32 Jacobi-PCG solves (each with its own
 * mesh, density, constraints, state, tolerance) in shared launches: one tile
 * assembly and one gather launch per (element kind, density kind, quadrature)
 * class and one vector launch per iteration for all slots.  No counterpart
 * in the reference (it is the serving loop's solver; the linear algebra is the
 * same as tatva_pcg_matfree_f32).  A slot's arithmetic does not depend on the
 * other slots, so its result is identical alone and in any mix. */
typedef struct tatva_batch tatva_batch;
int tatva_batch_create(int device, int32_t capacity, tatva_batch **out);
int tatva_batch_destroy(tatva_batch *batch);
/* Bytes of the per-slot workspace for a solve of n reduced dofs on mesh. */
int64_t tatva_batch_slot_workspace(tatva_mesh *mesh, int32_t ncomp, int64_t n);
/* (Re)starts slot: H(u) x = b from x (in place; updated by the iterations).
 * u (full FP32 state), dof_map, diag, b, x and the workspace must stay valid
 * until the slot is released or restarted.  Set-up kernels run on stream. */
int tatva_batch_start(tatva_batch *batch, int32_t slot, tatva_mesh *mesh,
                      const tatva_density *density, const float *u, const int32_t *dof_map,
                      int64_t n, const float *diag, const float *b, float *x, double tol,
                      int32_t maxiter, void *workspace, int64_t workspace_bytes, void *stream);
int tatva_batch_release(tatva_batch *batch, int32_t slot);
/* iters PCG iterations of every active slot (converged slots idle). */
int tatva_batch_iterate(tatva_batch *batch, int32_t iters, void *stream);
/* Synchronises stream and writes capacity stats records (reserved = done flag;
 * inactive slots are zero) and the number of active slots still iterating. */
int tatva_batch_poll(tatva_batch *batch, tatva_pcg_stats *stats, int32_t *n_running,
                     void *stream);

/* ------------------------------------------------------- vector kernels ---- */

/* out_full[free[i]] = in_red[i] for i < n_red (Lifter.lift, lifter/base.py:201-221;
 * the constrained entries of out_full are left as they are). */
int tatva_scatter_free(int64_t n_red, const int32_t *free_dofs, const double *in_red,
                       double *out_full, void *stream);
/* out_full[dofs[i]] = values[i] (or scalar value if values == NULL)
 * (Fixed.apply_lift, lifter/constraints.py:312-314). */
int tatva_set_dofs(int64_t n, const int32_t *dofs, const double *values, double value,
                   double *out_full, void *stream);
/* out_red[i] = in_full[free[i]] (Lifter.reduce, lifter/base.py:231-233). */
int tatva_gather_free(int64_t n_red, const int32_t *free_dofs, const double *in_full,
                      double *out_red, void *stream);
/* Linear part of a lift through a read map: out_full[d] = in_red[map[d]], or 0
 * where map[d] < 0.  With Fixed constraints only map is the free-dof index map;
 * a Periodic slave maps to its master's reduced index (Lifter.lift with the
 * constraint values at zero, lifter/base.py:201-221, constraints.py:214-216). */
int tatva_lift_map(int64_t n_full, const int32_t *map, const double *in_red, double *out_full,
                   void *stream);
/* out_red[j] = sum of in_full[list[k]], k in [ptr[j], ptr[j+1]) in ascending
 * order: the transpose of the lift (Lifter.reduce_adjoint, lifter/base.py:235-251,
 * Periodic.apply_transpose constraints.py:218-220); list holds every full dof
 * whose read map is j. */
int tatva_reduce_rows(int64_t n_red, const int32_t *ptr, const int32_t *list,
                      const double *in_full, double *out_red, void *stream);
/* u_full[free[i]] += alpha * du[i] with du FP32 or FP64. */
int tatva_axpy_free_f32(int64_t n_red, const int32_t *free_dofs, double alpha, const float *du,
                        double *u_full, void *stream);
int tatva_axpy_free(int64_t n_red, const int32_t *free_dofs, double alpha, const double *du,
                    double *u_full, void *stream);
/* y = x cast to FP32 / FP64, scaled by alpha. */
int tatva_cast_f64_f32(int64_t n, double alpha, const double *x, float *y, void *stream);
int tatva_cast_f32_f64(int64_t n, double alpha, const float *x, double *y, void *stream);
/* out[0] = x . y (deterministic reduction, FP64), the per-block partial sums in
 * the caller's workspace of tatva_dot_workspace(n) bytes (calls on different
 * streams need different workspaces). */
int64_t tatva_dot_workspace(int64_t n);
int tatva_dot(int64_t n, const double *x, const double *y, double *out, void *workspace,
              int64_t workspace_bytes, void *stream);
/* tatva_dot that allocates nothing and takes no workspace.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  The partial sums go to a
 * static device array of the library, 2048 doubles (16 KiB) on each device it
 * runs on, held until the process ends; the reduction grid is capped at 2048
 * blocks, so every n >= 1 fits.  Same value as tatva_dot, bit for bit.  All
 * calls on one device share that array: calls that overlap on one device (two
 * streams, two threads) race. */
int tatva_dot_noalloc(int64_t n, const double *x, const double *y, double *out, void *stream);

/* ------------------------------------------------ host-side setup (no GPU) -- */

typedef struct tatva_csr_host tatva_csr_host;

/* CSR sparsity pattern of the element connectivity with ndpn dofs per node:
 * every pair of element dofs, deduplicated, sorted (tatva/sparse/_extraction.py:
 * 37-102).  With dof_map (full dof -> reduced index or -1) the pattern is built
 * directly on the reduced dofs (= S[free][:, free], lifter/base.py:322-331). */
int tatva_pattern_from_mesh(const int32_t *elements, int64_t n_elements, int nne,
                            int64_t n_nodes, int ndpn, const int32_t *dof_map, int64_t n_rows,
                            tatva_csr_host **out);
/* CSR pattern of general element dof tables (the core of pattern_from_compound,
 * tatva/sparse/_extraction.py:125-245): every pair of valid (>= 0) dofs of one
 * element row of element_dofs (E x k, -1 = no dof), plus (d, d) for each
 * diag_dofs entry; deduplicated, columns sorted. */
int tatva_pattern_from_element_dofs(const int32_t *element_dofs, int64_t n_elements, int k,
                                    int64_t size, const int32_t *diag_dofs, int64_t n_diag,
                                    tatva_csr_host **out);
/* Row/column submatrix S[keep][:, keep] of a CSR pattern; keep_map: old index ->
 * new index or -1, monotone on kept indices (Lifter.reduce_sparsity,
 * lifter/base.py:322-331). */
int tatva_csr_submatrix(int64_t n, const int32_t *indptr, const int32_t *indices,
                        const int32_t *keep_map, int64_t n_new, tatva_csr_host **out);
int64_t tatva_csr_host_rows(const tatva_csr_host *csr);
int64_t tatva_csr_host_nnz(const tatva_csr_host *csr);
int tatva_csr_host_copy(const tatva_csr_host *csr, int32_t *indptr, int32_t *indices);
void tatva_csr_host_free(tatva_csr_host *csr);

/* Natural-order greedy distance-2 coloring of a CSR pattern (the in-repo
 * tatva/sparse/_coloring.py:27-48,136-153 algorithm used by ColoredMatrix.from_csr,
 * tatva/sparse/base.py:59-74).  Writes colors (n) and the number of colors. */
int tatva_distance2_colors(int64_t n, const int32_t *indptr, const int32_t *indices,
                           int32_t *colors, int32_t *n_colors);

/* Checks that colors (n, host) are a distance-2 coloring of a CSR pattern (host
 * arrays): no row has two columns of one color, which the colored
 * decompression of tatva_colored_hessian needs (ColoredMatrix.from_csr with
 * caller colors, tatva/sparse/base.py:59-74).  *bad_row = the first offending
 * row, or -1 when the coloring is valid. */
int tatva_check_distance2_colors(int64_t n, const int32_t *indptr, const int32_t *indices,
                                 const int32_t *colors, int64_t *bad_row);

/* ------------------------------------------------------------ measurement -- */

/* VALU issue-rate calibration: `blocks` workgroups of 256 threads, each thread
 * running eight independent FP32 chains for 32 x iters instructions of one
 * form (op 0: v_fma_f32 with two scalar operands, 1: v_fmac_f32 with vector
 * operands, 2: v_add_f32, 3: v_mul_f32, 4: v_fma_f32 with vector operands,
 * 5: v_fmac_f32 with one scalar operand), one float per thread written to out
 * (blocks x 256).  No counterpart in the reference: it
 * gives the roofline of the VALU-bound FP32 tile kernels a measured ceiling
 * (SQ_INSTS_VALU of this kernel over its duration), as a plain copy does for
 * the HBM traffic counters. */
int tatva_valu_calibration(int32_t op, int32_t blocks, int32_t iters, float *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* TATVA_AMD_H */
