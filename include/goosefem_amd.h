/*
 * goosefem_amd.h -- C ABI of the MI355X-native element-quadrature and assembly path.
 *
 * Plain pointers and sizes only.  Conventions:
 *   - every function returns 0 on success or a negative GF_E* code; gf_last_error() returns a
 *     thread-local message for the last failure (the reference throws std::runtime_error from
 *     GOOSEFEM_ASSERT / GOOSEFEM_CHECK, include/GooseFEM/config.h:65-98);
 *   - pointers named *_h are host memory, every other array pointer is device memory (HBM);
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream); device work is
 *     enqueued on it and the call returns without synchronising, unless stated otherwise;
 *   - index arrays handed in at set-up are int64 on the host (the reference uses size_t); the
 *     library keeps int32 copies on the device;
 *   - array layouts are the reference's row-major ones: elemvec [nelem, nne, ndim],
 *     qtensor [nelem, nip, tdim, tdim], qscalar [nelem, nip], elemmat [nelem, nne*ndim, nne*ndim],
 *     nodevec [nnode, ndim], dofval [ndof].
 * Each declaration names the reference interface it replaces (paths relative to the GooseFEM
 * repository root).
 */
#ifndef GOOSEFEM_AMD_H
#define GOOSEFEM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------------- */
#define GF_OK 0
#define GF_EINVAL -1   /* invalid argument / shape mismatch (GOOSEFEM_ASSERT) */
#define GF_EHIP -2     /* HIP runtime error */
#define GF_ENOMEM -3   /* device allocation failed */
#define GF_ESTATE -4   /* object not in the required state (e.g. matrix not assembled) */
#define GF_ENOTCONV -5 /* iterative solve reached its iteration cap */

/* element types */
#define GF_QUAD4 0        /* Element::Quad4::Quadrature, tdim = 2 */
#define GF_QUAD4_PLANAR 1 /* Element::Quad4::QuadraturePlanar, tdim = 3 (plane strain) */
#define GF_HEX8 2         /* Element::Hex8::Quadrature, tdim = 3 */

/* integration rules */
#define GF_GAUSS 0 /* Element::*::Gauss (ElementQuad4.h:36-93, ElementHex8.h:26-110) */
#define GF_NODAL 1 /* Element::*::Nodal (ElementQuad4.h:103-163, ElementHex8.h:116-200) */

/* arithmetic of element kernels */
#define GF_F64 0    /* FP64 throughout, reference expression order, no FMA contraction */
#define GF_F32 1    /* FP32 per-quadrature-point work (inputs/outputs FP32 where typed void*) */
#define GF_BF16X3 2 /* stiffness only: BF16 hi/lo split operands on MFMA, FP32 accumulation */
#define GF_F64_FAST 3 /* explicit step only: FP64 throughout with fused multiply-add, one division per
                       * Gauss point and the closed-form rectangle body on Regular coordinates;
                       * within 1e-14 of GF_F64, not bit for bit */

typedef struct gf_vector_s* gf_vector;
typedef struct gf_quad_s* gf_quad;
typedef struct gf_diag_s* gf_diag;
typedef struct gf_matrix_s* gf_matrix;
typedef struct gf_explicit_s* gf_explicit;
typedef struct gf_pcg_s* gf_pcg;
typedef struct gf_vecpart_s* gf_vecpart;
typedef struct gf_diagpart_s* gf_diagpart;
typedef struct gf_matpart_s* gf_matpart;
typedef struct gf_plastic_s* gf_plastic;
typedef struct gf_axi_s* gf_axi;
typedef struct gf_mattyings_s* gf_mattyings;

const char* gf_last_error(void);
int gf_version(int* major, int* minor, int* patch);
/* Synchronise a stream (host-side wait).  Convenience for FFI callers. */
int gf_stream_sync(void* stream);

/* Device memory for FFI callers without HIP headers (the C++ shim uses these). */
int gf_device_malloc(void** ptr, int64_t bytes);
int gf_device_free(void* ptr);
int gf_memcpy_h2d(void* dst, const void* src_h, int64_t bytes);
int gf_memcpy_d2h(void* dst_h, const void* src, int64_t bytes);
int gf_memcpy_d2d(void* dst, const void* src, int64_t bytes);
/* Streaming probes over n doubles (n even, 16-byte aligned arrays of n doubles each): the measured
 * HBM ceilings that roofline fractions are reported against beside the datasheet peak.  mode 0:
 * c = a; 1: c = a with non-temporal accesses; 2: c = a + b (non-temporal); 3: read a only (c is
 * written only for data that the probe never uses).  No counterpart in the reference (measurement
 * utility). */
int gf_stream_probe(int mode, const double* a, const double* b, double* c, int64_t n, void* stream);

/* ================================================================================================
 * Mesh bookkeeping (host, set-up only).  Integer outputs are bit-exact with the reference.
 * Two-call pattern for variable-size outputs: pass NULL to get the size.
 * ============================================================================================== */

/* Mesh::Quad4::Regular(nelx, nely, h).coor() / .conn()   (include/GooseFEM/MeshQuad4.h:43-122)
 * coor_h [(nelx+1)(nely+1), 2], conn_h [nelx*nely, 4] */
int gf_mesh_quad4_regular(int64_t nelx, int64_t nely, double h, double* coor_h, int64_t* conn_h);
/* Mesh::Hex8::Regular(nelx, nely, nelz, h).coor() / .conn()   (include/GooseFEM/MeshHex8.h:41-131) */
int gf_mesh_hex8_regular(int64_t nelx, int64_t nely, int64_t nelz, double h, double* coor_h,
                         int64_t* conn_h);
/* Named node sets of the Regular meshes, e.g. "nodesBottomEdge", "nodesTopOpenEdge",
 * "nodesFrontFace", "nodesBackTopRightCorner", "nodesOrigin" (MeshQuad4.h:128-187,
 * MeshHex8.h:133-597, Mesh.h:264-330).  nelz <= 0 selects the 2-D mesh.  *n receives the count. */
int gf_mesh_nodeset(int64_t nelx, int64_t nely, int64_t nelz, const char* name, int64_t* out_h,
                    int64_t* n);
/* RegularBase::nodesPeriodic()  (Mesh.h:511-540 for 2-D, Mesh.h:1577-1692 for 3-D).
 * out_h [nties, 2] (independent, dependent); *nties receives the count. */
int gf_mesh_nodes_periodic(int64_t nelx, int64_t nely, int64_t nelz, int64_t* out_h, int64_t* nties);
/* Mesh::Quad4::FineLayer(nelx, nely, h, nfine)  (include/GooseFEM/MeshQuad4.h:200-1135): element
 * layers finest around the middle, coarsened by 3 in x towards the bottom and top.  sizes (any
 * output may be NULL) receives {nelem, nnode, nelx(), nely()} (nelx() the finest layer's element
 * count, nely() the height in fine elements, MeshQuad4.h:571-579); coor_h [nnode, 2], conn_h
 * [nelem, 4]. */
int gf_mesh_quad4_finelayer(int64_t nelx, int64_t nely, double h, int64_t nfine, int64_t* sizes, double* coor_h,
                            int64_t* conn_h);
/* Named sets of that mesh: "nodesBottomEdge" ... "nodesRightOpenEdge", the corners,
 * "nodesOrigin", "elementsMiddleLayer" (MeshQuad4.h:287-292, 747-880) and "nodesPeriodic" (pairs,
 * flattened [nties * 2], the rule of Mesh.h:511-540).  *n receives the count. */
int gf_mesh_quad4_finelayer_set(int64_t nelx, int64_t nely, double h, int64_t nfine, const char* name,
                                int64_t* out_h, int64_t* n);
/* FineLayer::roll(n)  (include/GooseFEM/MeshQuad4.h:539-564): the element map (out_h [nelem],
 * new_elemvar = elemvar[out_h]) that shifts every element layer periodically in x by n elements
 * of the coarsest layer. */
int gf_mesh_quad4_finelayer_roll(int64_t nelx, int64_t nely, double h, int64_t nfine, int64_t n, int64_t* out_h);
/* Mesh::Hex8::FineLayer(nelx, nely, nelz, h, nfine)  (include/GooseFEM/MeshHex8.h:613-2296):
 * element layers finest around the middle, coarsened by 3 in x or in z towards the bottom and top.
 * sizes (any output may be NULL) receives {nelem, nnode, nelx(), nely(), nelz()} (nelx(), nelz()
 * the finest layer's element counts, nely() the height in fine elements, MeshHex8.h:874-886);
 * coor_h [nnode, 3], conn_h [nelem, 8]. */
int gf_mesh_hex8_finelayer(int64_t nelx, int64_t nely, int64_t nelz, double h, int64_t nfine, int64_t* sizes,
                           double* coor_h, int64_t* conn_h);
/* Named sets of that mesh: the faces ("nodesFront" ...), face interiors ("nodesFrontFace" ...),
 * edges and open edges ("nodesFrontBottomEdge", "nodesFrontBottomOpenEdge" ..., the words in any
 * order), corners, "nodesOrigin", "elementsMiddleLayer" (MeshHex8.h:855-870, 1331-2275) and
 * "nodesPeriodic" (pairs, flattened [nties * 2], the rule of Mesh.h:1577-1692).  *n receives the
 * count. */
int gf_mesh_hex8_finelayer_set(int64_t nelx, int64_t nely, int64_t nelz, double h, int64_t nfine, const char* name,
                               int64_t* out_h, int64_t* n);
/* Mesh::dofs(nnode, ndim)  (Mesh.h:93-96) */
int gf_mesh_dofs(int64_t nnode, int ndim, int64_t* dofs_h);
/* RegularBase::dofsPeriodic()  (Mesh.h:291-304): dependent nodes take their independent node's
 * dofs, then Mesh::renumber (Mesh.h:114-175). */
int gf_mesh_dofs_periodic(int64_t nnode, int ndim, const int64_t* ties_h, int64_t nties,
                          int64_t* dofs_h);
/* Mesh::renumber  (Mesh.h:114-175) */
int gf_mesh_renumber(const int64_t* in_h, int64_t n, int64_t* out_h);
/* Mesh::coordination  (Mesh.h:2462-2477): out_h [max(conn)+1] */
int gf_mesh_coordination(const int64_t* conn_h, int64_t n, int64_t nvalues, int64_t* out_h);
/* Mesh::elem2node(conn[, dofs])  (Mesh.h:2485-2551) as CSR: ptr_h [nnode+1], idx_h [nnz].
 * dofs_h may be NULL (non-periodic).  *nnz receives the entry count. */
int gf_mesh_elem2node(const int64_t* conn_h, int64_t nelem, int nne, const int64_t* dofs_h,
                      int64_t nnode, int ndim, int64_t* ptr_h, int64_t* idx_h, int64_t* nnz);
/* Mesh::nodaltyings(dofs)  (Mesh.h:2839-2861) as CSR. */
int gf_mesh_nodaltyings(const int64_t* dofs_h, int64_t nnode, int ndim, int64_t* ptr_h,
                        int64_t* idx_h, int64_t* nnz);

/* Tyings::Control(coor, dofs)  (include/GooseFEM/TyingsPeriodic.h:361-392).
 * coor_out_h [nnode+ndim, ndim], dofs_out_h [nnode+ndim, ndim], control_dofs_h [ndim, ndim],
 * control_nodes_h [ndim]. */
int gf_tyings_control(const double* coor_h, const int64_t* dofs_h, int64_t nnode, int ndim,
                      double* coor_out_h, int64_t* dofs_out_h, int64_t* control_dofs_h,
                      int64_t* control_nodes_h);
/* Tyings::Periodic(coor, dofs, control, nodal_tyings, iip)  (TyingsPeriodic.h:82-121):
 * reordered dofs_out_h [nnode, ndim] ([u, p, d] numbering, Mesh::Reorder Mesh.h:2384-2454),
 * control_out_h [ndim, ndim], sizes_h = {nnu, nnp, nni, nnd}. */
int gf_tyings_periodic(const int64_t* dofs_h, int64_t nnode, int ndim, const int64_t* control_h,
                       const int64_t* ties_h, int64_t nties, const int64_t* iip_h, int64_t niip,
                       int64_t* dofs_out_h, int64_t* control_out_h, int64_t* sizes_h);
/* Tyings::Periodic::Cdi() / Cdu() / Cdp()  (TyingsPeriodic.h:238-343) as CSR with sorted columns,
 * duplicates summed.  which: 0 = Cdi, 1 = Cdu, 2 = Cdp.  dofs_h / control_h are the reordered
 * ones.  rowptr_h [nties*ndim + 1]; col_h/val_h may be NULL to query *nnz. */
int gf_tyings_matrix(int which, const double* coor_h, const int64_t* dofs_h, int ndim,
                     const int64_t* control_h, const int64_t* ties_h, int64_t nties, int64_t nnu,
                     int64_t* rowptr_h, int64_t* col_h, double* val_h, int64_t* nnz);

/* Portable synthetic input generator (SplitMix64 -> Box-Muller, host FP64).  No reference
 * counterpart; used to build identical inputs on every machine. */
int gf_randn(uint64_t seed, int64_t n, double mean, double stddev, double* out_h);

/* ================================================================================================
 * Vector: storage conversions  (include/GooseFEM/Vector.h)
 * Scatter-adds are deterministic node/DOF-centric gathers over precomputed CSR lists sorted in the
 * reference's serial loop order, so FP64 sums are bit-identical to the serial reference loop.
 * ============================================================================================== */

/* Vector(conn, dofs)  (Vector.h:33-47); ndof = max(dofs) + 1 */
int gf_vector_create(gf_vector* out, const int64_t* conn_h, int64_t nelem, int nne,
                     const int64_t* dofs_h, int64_t nnode, int ndim);
int gf_vector_destroy(gf_vector v);
int gf_vector_sizes(gf_vector v, int64_t* nelem, int* nne, int64_t* nnode, int* ndim, int64_t* ndof);
/* device copies of conn / dofs (int32), valid for the lifetime of v */
int gf_vector_device_indices(gf_vector v, const int32_t** conn, const int32_t** dofs);
/* AsElement(nodevec) / asElement(nodevec, elemvec)  (Vector.h:193-211, :680-697) */
int gf_vector_as_element_nodevec(gf_vector v, const double* nodevec, double* elemvec, void* stream);
/* AsElement(dofval)  (Vector.h:193-211, :658-675) */
int gf_vector_as_element_dofval(gf_vector v, const double* dofval, double* elemvec, void* stream);
/* AsDofs(nodevec)  (Vector.h:141-159, :568-585): overwrite, last node in loop order wins */
int gf_vector_as_dofs_nodevec(gf_vector v, const double* nodevec, double* dofval, void* stream);
/* AsDofs(elemvec)  (Vector.h:141-159, :590-609): overwrite, last (e, m) in loop order wins */
int gf_vector_as_dofs_elemvec(gf_vector v, const double* elemvec, double* dofval, void* stream);
/* AsNode(dofval)  (Vector.h:167-185, :614-629) */
int gf_vector_as_node_dofval(gf_vector v, const double* dofval, double* nodevec, void* stream);
/* AsNode(elemvec)  (Vector.h:167-185, :634-653) */
int gf_vector_as_node_elemvec(gf_vector v, const double* elemvec, double* nodevec, void* stream);
/* AssembleDofs(nodevec)  (Vector.h:219-237, :702-718) */
int gf_vector_assemble_dofs_nodevec(gf_vector v, const double* nodevec, double* dofval, void* stream);
/* AssembleDofs(elemvec)  (Vector.h:219-237, :723-741).  elemvec is FP64 (dtype GF_F64) or FP32
 * (GF_F32, summed in FP64). */
int gf_vector_assemble_dofs_elemvec(gf_vector v, const void* elemvec, int dtype, double* dofval,
                                    void* stream);
/* AssembleNode(elemvec)  (Vector.h:245-263, :746-757) = AsNode(AssembleDofs(elemvec)) */
int gf_vector_assemble_node_elemvec(gf_vector v, const void* elemvec, int dtype, double* nodevec,
                                    void* stream);

/* ================================================================================================
 * Quadrature  (include/GooseFEM/Element.h:545-868, ElementQuad4.h, ElementQuad4Planar.h,
 * ElementHex8.h).  Geometry (GradN, dV) is computed on the device at construction.
 * ============================================================================================== */

/* Quadrature(x) / Quadrature(x, xi, w) / QuadraturePlanar(x, thick)  (ElementQuad4.h:255-318,
 * ElementQuad4Planar.h:58-123, ElementHex8.h:233-317).  x: device elemvec of coordinates.
 * xi_h [nip, ndim] and w_h [nip] may be NULL to select the rule `rule`.  The Hex8 geometry uses
 * the chain rule dNdx = J^-1 dN/dxi (the generic reference loop, Element.h:1220, indexes
 * dN/dxi(m,i) for dN/dxi(m,j); both agree to round-off on every Regular mesh). */
int gf_quad_create(gf_quad* out, int etype, const double* x, int64_t nelem, int rule,
                   const double* xi_h, const double* w_h, int nip, double thick, void* stream);
int gf_quad_destroy(gf_quad q);
int gf_quad_sizes(gf_quad q, int64_t* nelem, int* nne, int* ndim, int* tdim, int* nip);
/* update_x(x)  (Element.h:563-569) */
int gf_quad_update_x(gf_quad q, const double* x, void* stream);
/* GradN() [nelem, nip, nne, ndim] and dV() [nelem, nip]  (Element.h:575-587): pointers to the
 * internal device arrays (aliases, like the reference's const references). */
int gf_quad_gradN(gf_quad q, const double** dNx);
int gf_quad_dV(gf_quad q, const double** dV);
/* InterpQuad_vector(elemvec) -> qvector [nelem, nip, n]  (Element.h:597-616) */
int gf_quad_interp_quad_vector(gf_quad q, const double* elemvec, int n, double* qvector, void* stream);
/* GradN_vector  (Element.h:634-652): qtensor(i,j) = dNx(m,i) u(m,j).  dtype GF_F64 or GF_F32 for
 * both elemvec and qtensor. */
int gf_quad_gradN_vector(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream);
/* GradN_vector_T  (Element.h:666-684) */
int gf_quad_gradN_vector_T(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream);
/* SymGradN_vector  (Element.h:699-717) */
int gf_quad_symgradN_vector(gf_quad q, const void* elemvec, void* qtensor, int dtype, void* stream);
/* Int_N_vector_dV(qvector [nelem, nip, n]) -> elemvec [nelem, nne, n]  (Element.h:731-750) */
int gf_quad_int_N_vector_dV(gf_quad q, const double* qvector, int n, double* elemvec, void* stream);
/* Int_N_scalar_NT_dV(qscalar) -> elemmat  (Element.h:770-788) */
int gf_quad_int_N_scalar_NT_dV(gf_quad q, const double* qscalar, double* elemmat, void* stream);
/* Int_gradN_dot_tensor2_dV(qtensor) -> elemvec  (Element.h:807-825); dtype as above */
int gf_quad_int_gradN_dot_tensor2_dV(gf_quad q, const void* qtensor, void* elemvec, int dtype,
                                     void* stream);
/* Int_gradN_dot_tensor4_dot_gradNT_dV(qtensor4 [nelem, nip, tdim^4]) -> elemmat [FP64]
 * (Element.h:849-868).  precision GF_F64 (reference order) or GF_BF16X3 (MFMA, Hex8). */
int gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV(gf_quad q, const double* qtensor4, double* elemmat,
                                                int precision, void* stream);
/* The same product (Element.h:849-868) at GF_BF16X3 with FP32 qtensor4 in and FP32 elemmat out
 * (Hex8, 8 points): the kernel's operands are FP32 before the hi/lo split and it accumulates in
 * FP32, so FP32 input and output carry every digit it uses, in 57% of the FP64 interface's bytes. */
int gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV_f32(gf_quad q, const float* qtensor4, float* elemmat,
                                                    void* stream);

/* Batched element matrix-vector product out[e] = elemmat[e] . elemvec[e] (K = nne*ndim), the
 * element part of the matrix-free product sum_e K_e u_e; assemble with
 * gf_vector_assemble_dofs_elemvec.  No single reference function: it is Matrix::Dot
 * (Matrix.h:804-815) without forming the sparse matrix. */
int gf_elemmat_dot(int64_t nelem, int K, const double* elemmat, const double* elemvec, double* out,
                   void* stream);
/* The same with an FP32 elemmat (each entry widened to FP64 before its product; FP64 sums) */
int gf_elemmat_dot_f32(int64_t nelem, int K, const float* elemmat, const double* elemvec, double* out,
                       void* stream);

/* Material: linear elastic law of the callers (GMat* in docs/examples; not in the reference):
 * sigma = K tr(eps) I + 2 G dev(eps), per quadrature point; K_q/G_q [nelem*nip] FP64, or FP32
 * for dtype GF_F32 (then eps/sig are FP32 too). */
int gf_material_elastic_stress(int64_t nq, int tdim, const void* eps, const void* K_q,
                               const void* G_q, void* sig, int dtype, void* stream);
/* Tangent C_ijkl = K I(x)I + 2G (I_sym - I(x)I/3), FP64 [nq, tdim^4] */
int gf_material_elastic_tangent(int64_t nq, int tdim, const double* K_q, const double* G_q, double* C,
                                void* stream);
/* The same tangent rounded to FP32 (the input of gf_quad_int_gradN_dot_tensor4_dot_gradNT_dV_f32) */
int gf_material_elastic_tangent_f32(int64_t nq, int tdim, const double* K_q, const double* G_q, float* C,
                                    void* stream);

/* ================================================================================================
 * MatrixPartitionedTyings + MatrixPartitionedTyingsSolver for any mesh and element matrices
 * (include/GooseFEM/MatrixPartitionedTyings.h:34-675).  K is assembled as one CSR over the tied
 * numbering [u | p | d] (a gf_matrix over v, whose dofs are Tyings::Periodic's dofs()); the
 * condensed operator K'_uu = K_uu + K_ud C_du + C_ud K_du + C_ud K_dd C_du (:661-663) is applied
 * matrix-free (expand x_d = C_du x_u + C_dp x_p, one CSR product, fold y_u + C_ud y_d) inside an
 * FP64 Jacobi-preconditioned CG instead of an LDLT factorisation, so K'_uu must be symmetric
 * positive definite.  C_du [nnd, nnu] and C_dp [nnd, nnp] are host CSR (int64 rowptr / col, FP64
 * values), as Tyings::Periodic::Cdu() / Cdp() (TyingsPeriodic.h:244-256).
 * ============================================================================================== */
int gf_mattyings_create(gf_mattyings* out, gf_vector v, int64_t nnu, int64_t nnp, int64_t nnd, const int64_t* du_rowptr,
                        const int64_t* du_col, const double* du_val, const int64_t* dp_rowptr, const int64_t* dp_col,
                        const double* dp_val);
int gf_mattyings_destroy(gf_mattyings A);
/* the unpartitioned CSR (Todense via gf_matrix_todense) */
int gf_mattyings_matrix(gf_mattyings A, gf_matrix* K);
/* assemble(elemmat)  (MatrixPartitionedTyings.h:210-305): K, and the Jacobi diagonal of K' */
int gf_mattyings_assemble(gf_mattyings A, const double* elemmat, void* stream);
/* MatrixPartitionedTyingsSolver::solve(A, b, x)  (:540-564) on nodevecs: B_u = b_u + C_ud b_d,
 * X_u = K'^-1 (B_u - K'_up X_p), X_d = C_du X_u + C_dp X_p; writes x's u and d entries, keeps its
 * p entries.  Stops at ||r|| <= tol ||rhs||; GF_ENOTCONV if the true residual exceeds 10 tol. */
int gf_mattyings_solve(gf_mattyings A, const double* b, double* x, double tol, int64_t maxiter, int64_t* iterations,
                       double* relres, void* stream);

/* y += A x for a device CSR A (int32 rowptr [nrows + 1], col; FP64 values), each row summed in
 * entry order onto y's value.  No counterpart in the reference: it applies the transposed tying
 * matrix in VectorPartitionedTyings::asDofs_i (VectorPartitionedTyings.h:224-251). */
int gf_csr_addmv(int64_t nrows, const int32_t* rowptr, const int32_t* col, const double* val, const double* x, double* y,
                 void* stream);

/* ================================================================================================
 * Element::Quad4::QuadratureAxisymmetric  (include/GooseFEM/ElementQuad4Axisymmetric.h:33-467)
 * Nodal coordinates (z, r); tensors 3 x 3 in (r, theta, z) order.  The geometry is kept as
 * G [nelem, nip, 4, 3] = (dN/dr, N / r_q, dN/dz) per node, the five non-zero components of the
 * reference's B (:401-438), and dV [nelem, nip] = w J 2 pi r_q.  Results equal the reference's bit
 * for bit (same FP64 expression order).  x [nelem, 4, 2] device; xi_h [nip, 2], w_h [nip] host
 * (nip <= 16).
 * ============================================================================================== */
int gf_axi_create(gf_axi* out, int64_t nelem, int nip, const double* x, const double* xi_h, const double* w_h,
                  void* stream);
int gf_axi_destroy(gf_axi m);
/* update_x (Element.h:563-587 for this element) */
int gf_axi_update_x(gf_axi m, const double* x, void* stream);
/* device pointers to G and dV (dV(), :446); either output may be NULL */
int gf_axi_geometry(gf_axi m, const double** G, const double** dV);
/* mode 0: gradN_vector (:140-169); 1: gradN_vector_T (:171-200); 2: symGradN_vector (:202-233);
 * elemvec [nelem, 4, 2] -> qtensor [nelem, nip, 3, 3] */
int gf_axi_grad(gf_axi m, int mode, const double* elemvec, double* qtensor, void* stream);
/* int_N_scalar_NT_dV (:235-266): qscalar [nelem, nip] -> elemmat [nelem, 8, 8] */
int gf_axi_int_N_scalar_NT_dV(gf_axi m, const double* qscalar, double* elemmat, void* stream);
/* int_gradN_dot_tensor2_dV (:268-296): qtensor [nelem, nip, 3, 3] -> elemvec [nelem, 4, 2] */
int gf_axi_int_gradN_dot_tensor2_dV(gf_axi m, const double* qtensor, double* elemvec, void* stream);
/* int_gradN_dot_tensor4_dot_gradNT_dV (:298-398): qtensor [nelem, nip, 3, 3, 3, 3] -> elemmat [nelem, 8, 8] */
int gf_axi_int_gradN_dot_tensor4_dot_gradNT_dV(gf_axi m, const double* qtensor, double* elemmat, void* stream);

/* Material: J2 elasto-plasticity with linear isotropic hardening, the law the statics examples call
 * as GMatElastoPlastic.Cartesian3d.LinearHardening2d(K, G, sigy0, H)
 * (docs/examples/statics/periodic_elastoplastic.py:71-104; external package, not in the
 * reference).  Radial return and consistent tangent of de Souza Neto et al. (2008) Box 7.3 and
 * eq. 7.120.  All arrays are device FP64: K, G, sigy0, H [nq] (copied at create); Eps, Sig, Epsp
 * [nq, 3, 3]; C [nq, 3, 3, 3, 3]; epsp [nq].  The state starts at zero plastic strain. */
int gf_plastic_create(gf_plastic* out, int64_t nq, const double* K, const double* G, const double* sigy0,
                      const double* H, void* stream);
int gf_plastic_destroy(gf_plastic m);
/* refresh(): Sig (and C unless NULL) from Eps and the committed state; updates the trial state
 * Epsp, epsp */
int gf_plastic_refresh(gf_plastic m, const double* Eps, double* Sig, double* C, void* stream);
/* increment(): commit the state (Epsp_t = Epsp, epsp_t = epsp) */
int gf_plastic_increment(gf_plastic m, void* stream);
/* One pass of the statics examples' Newton iteration on a Quad4 QuadraturePlanar with four points
 * (periodic_elastoplastic.py:87-104), fused: Eps = symGradN_vector(elemvec); refresh();
 * fe = int_gradN_dot_tensor2_dV(Sig) [nelem, 4, 2]; Ke = int_gradN_dot_tensor4_dot_gradNT_dV(C)
 * [nelem, 8, 8]; Sig [nelem, 4, 3, 3] is written only if not NULL.  Eps and C never reach memory;
 * the results equal the unfused calls' bit for bit.  fe, Ke and Sig must be 16-byte aligned
 * (GF_EINVAL otherwise).  No counterpart in the reference (which composes the three calls). */
int gf_plastic_newton_q4(gf_plastic m, gf_quad q, const double* elemvec, double* fe, double* Ke, double* Sig,
                         void* stream);
/* device pointers to the state arrays; any output may be NULL */
int gf_plastic_state(gf_plastic m, const double** Epsp, const double** epsp, const double** Epsp_t,
                     const double** epsp_t);

/* ================================================================================================
 * MatrixDiagonal  (include/GooseFEM/MatrixDiagonal.h)
 * ============================================================================================== */

/* MatrixDiagonal(conn, dofs)  (MatrixDiagonal.h:132-147): uses v's conn/dofs */
int gf_diag_create(gf_diag* out, gf_vector v);
int gf_diag_destroy(gf_diag d);
/* assemble(elemmat): sums diagonal entries only, serial order  (MatrixDiagonal.h:150-167) */
int gf_diag_assemble(gf_diag d, const double* elemmat, void* stream);
/* set(A)  (MatrixDiagonal.h:185-190) */
int gf_diag_set(gf_diag d, const double* A, void* stream);
/* data()  (MatrixDiagonal.h:206-209): device pointer to A [ndof] */
int gf_diag_data(gf_diag d, const double** A);
/* inverse (factorize, MatrixDiagonal.h:268-280), computed lazily; device pointer [ndof] */
int gf_diag_inverse(gf_diag d, const double** inv, void* stream);
/* Dot(x)  (MatrixDiagonal.h:212-233) */
int gf_diag_dot_dofval(gf_diag d, const double* x, double* b, void* stream);
int gf_diag_dot_nodevec(gf_diag d, const double* x, double* b, void* stream);
/* Solve(b)  (MatrixDiagonal.h:236-259) */
int gf_diag_solve_dofval(gf_diag d, const double* b, double* x, void* stream);
int gf_diag_solve_nodevec(gf_diag d, const double* b, double* x, void* stream);

/* ================================================================================================
 * Dirichlet partitioning: VectorPartitioned (include/GooseFEM/VectorPartitioned.h) and
 * MatrixDiagonalPartitioned (include/GooseFEM/MatrixDiagonalPartitioned.h).  The DOFs split into
 * unknown iiu = setdiff1d(dofs, iip) (ascending) and prescribed iip (caller's order); part =
 * Mesh::Reorder({iiu, iip}).apply(dofs) (Mesh.h:2384-2454).  All results are bit-exact (gathers,
 * scatters and per-DOF products).  kind: 0 = dofval [ndof], 1 = nodevec [nnode, ndim],
 * 2 = elemvec [nelem, nne, ndim].
 * ============================================================================================== */

/* VectorPartitioned(conn, dofs, iip)  (VectorPartitioned.h:72-87); uses v's conn/dofs */
int gf_vecpart_create(gf_vecpart* out, gf_vector v, const int64_t* iip_h, int64_t nnp);
int gf_vecpart_destroy(gf_vecpart p);
/* nnu(), nnp()  (VectorPartitioned.h:92-103) */
int gf_vecpart_sizes(gf_vecpart p, int64_t* nnu, int64_t* nnp);
/* iiu(), iip() and the reordered DOFs part [nnode, ndim] as host copies (VectorPartitioned.h:108-119;
 * dofs_is_u / dofs_is_p, :126-161, are part < nnu / part >= nnu); any output may be NULL */
int gf_vecpart_index(gf_vecpart p, int64_t* iiu_h, int64_t* iip_h, int64_t* part_h);
/* AsDofs_u / AsDofs_p of a dofval, nodevec or elemvec (VectorPartitioned.h:428-661; nodevec and
 * elemvec overwrite in loop order: each DOF gets its last source); either output may be NULL */
int gf_vecpart_as_dofs(gf_vecpart p, int kind, const double* in, double* dofval_u, double* dofval_p, void* stream);
/* DofsFromParitioned / NodeFromPartitioned / ElementFromPartitioned (VectorPartitioned.h:276-424) */
int gf_vecpart_from_partitioned(gf_vecpart p, int kind, const double* dofval_u, const double* dofval_p, double* out,
                                void* stream);
/* copy_u (which 0) / copy_p (which 1): nodevec_dest[is_u or is_p] = nodevec_src  (VectorPartitioned.h:198-263) */
int gf_vecpart_copy(gf_vecpart p, int which, const double* nodevec_src, double* nodevec_dest, void* stream);

/* MatrixDiagonalPartitioned(conn, dofs, iip)  (MatrixDiagonalPartitioned.h:42-68) over p.  Its
 * unpartitioned diagonal is a gf_diag (gf_diagpart_diag): assemble (:73-98, the same per-DOF sum
 * in the same order), data (:142-156), Dot (:189-224) and the factorisation (:440-452,
 * inv = 1 / A) are the gf_diag_* calls on it. */
int gf_diagpart_create(gf_diagpart* out, gf_vecpart p);
int gf_diagpart_destroy(gf_diagpart d);
int gf_diagpart_diag(gf_diagpart d, gf_diag* out);
/* set(A)  (MatrixDiagonalPartitioned.h:121-135): A [ndof]; entries of DOF numbers that occur in
 * neither iiu nor iip are not taken (they read 0 in data()) */
int gf_diagpart_set(gf_diagpart d, const double* A, void* stream);
/* data_uu(), data_pp()  (MatrixDiagonalPartitioned.h:163-175); either output may be NULL */
int gf_diagpart_data_part(gf_diagpart d, double* A_uu, double* A_pp, void* stream);
/* Dot_u (which 0): b_u = A_uu x_u;  Dot_p / Reaction_p (which 1): b_p = A_pp x_p
 * (MatrixDiagonalPartitioned.h:248-300, :405-420) */
int gf_diagpart_dot_part(gf_diagpart d, int which, const double* x, double* b, void* stream);
/* Solve_u: x_u = A_uu^-1 b_u  (MatrixDiagonalPartitioned.h:356-372) */
int gf_diagpart_solve_u(gf_diagpart d, const double* b_u, double* x_u, void* stream);
/* Solve (op 0; writes only the unknown entries of out, :306-337) or Reaction (op 1; writes only
 * the prescribed entries, :378-403) on a dofval (kind 0) or nodevec (kind 1) */
int gf_diagpart_apply_sel(gf_diagpart d, int op, int kind, const double* in, double* out, void* stream);

/* ================================================================================================
 * Matrix: sparse (CSR, sorted columns)  (include/GooseFEM/Matrix.h:653-860)
 * Assembly is a deterministic gather: every CSR entry sums its element contributions in the
 * reference's triplet order (ascending e), so FP64 values match the reference's duplicate sum.
 * ============================================================================================== */

/* Matrix(conn, dofs)  (Matrix.h:679-693): builds the sparsity pattern and gather lists, and (rows
 * of at most 256 entries, elements of at most 32 DOFs without a repeated DOF) a device map of one
 * byte per element-matrix entry: the entry's position in its CSR row (nelem * K * K bytes). */
int gf_matrix_create(gf_matrix* out, gf_vector v);
int gf_matrix_destroy(gf_matrix A);
int gf_matrix_sizes(gf_matrix A, int64_t* nrows, int64_t* nnz);
/* assemble(elemmat)  (Matrix.h:704-727) */
int gf_matrix_assemble(gf_matrix A, const double* elemmat, void* stream);
/* assemble(elemmat) from FP32 element matrices: each entry widened to FP64, summed in the same
 * order as gf_matrix_assemble (Matrix.h:704-727) */
int gf_matrix_assemble_f32(gf_matrix A, const float* elemmat, void* stream);
/* device CSR arrays (rowptr [nrows+1], col [nnz] int32; val [nnz] FP64) */
int gf_matrix_csr(gf_matrix A, const int32_t** rowptr, const int32_t** col, const double** val);
/* Dot(x)  (Matrix.h:804-815): b = A x on dofval; nodevec variant via AsDofs / AsNode */
int gf_matrix_dot_dofval(gf_matrix A, const double* x, double* b, void* stream);
int gf_matrix_dot_nodevec(gf_matrix A, const double* x, double* b, void* stream);
/* gf_matrix_dot_nodevec without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The dofval staging of x and A x lives in a static device array of the
 * library, 2 x 2^20 doubles (16 MiB) per device: matrices above 2^20 rows fail with GF_EINVAL.
 * Calls that overlap on one GPU share that array and race; the stream must belong to the current
 * device.  Results are bit-identical to gf_matrix_dot_nodevec, which allocates and synchronises. */
int gf_matrix_dot_nodevec_noalloc(gf_matrix A, const double* x, double* b, void* stream);
/* Todense()  (Matrix.h:338-355, :790-802): out [nrows, nrows] */
int gf_matrix_todense(gf_matrix A, double* out, void* stream);

/* ================================================================================================
 * MatrixPartitioned / MatrixPartitionedSolver  (include/GooseFEM/MatrixPartitioned.h:34-640)
 * The blocks A_uu, A_up, A_pu, A_pp are CSR in the partitioned numbering (iiu then iip, columns
 * ascending).  Every block entry is a copy of the unpartitioned CSR entry (gf_matrix over the same
 * Vector) it comes from, so it holds the same element contributions summed in the same
 * (e, m, i, n, j) order as setFromTriplets (:156-210): assembly is bit-exact.  Products run one
 * sum per row over ascending columns of the first block, then of the second: the order Eigen
 * evaluates "A_uu X_u + A_up X_p" for a column-major matrix (dst = A X_u; dst += B X_p).
 * Solve factorises nothing: x_u = A_uu^-1 (b_u - A_up x_p) by a Jacobi-preconditioned CG on A_uu
 * (FP64, deterministic reductions), so A_uu must be symmetric positive definite (the reference's
 * SimplicialLDLT also takes indefinite symmetric matrices).  set() / add() of dense sub-matrices
 * are not provided (the sparsity pattern is the mesh's).
 * ============================================================================================== */

/* MatrixPartitioned(conn, dofs, iip)  (MatrixPartitioned.h:87-121) over p's Vector and iip.  Holds
 * an unpartitioned gf_matrix (gf_matpart_matrix: Todense, :324-350, is gf_matrix_todense on it) and
 * the four blocks; the pattern is built on the host once. */
int gf_matpart_create(gf_matpart* out, gf_vecpart p);
int gf_matpart_destroy(gf_matpart A);
/* nnz of the blocks uu, up, pu, pp (data_uu() ... data_pp(), :126-153) */
int gf_matpart_sizes(gf_matpart A, int64_t* nnz4);
int gf_matpart_matrix(gf_matpart A, gf_matrix* full);
/* device CSR of block 0 (uu), 1 (up), 2 (pu) or 3 (pp); rows nnu or nnp; outputs may be NULL */
int gf_matpart_csr(gf_matpart A, int block, const int32_t** rowptr, const int32_t** col, const double** val);
/* assemble(elemmat)  (MatrixPartitioned.h:156-210); the FP32 form widens each entry to FP64 */
int gf_matpart_assemble(gf_matpart A, const double* elemmat, void* stream);
int gf_matpart_assemble_f32(gf_matpart A, const float* elemmat, void* stream);
/* which 0: b = A_uu x_u + A_up x_p  (B_u of :378-396);  which 1: b = A_pu x_u + A_pp x_p
 * (Reaction_p, :436-447) */
int gf_matpart_dot_part(gf_matpart A, int which, const double* x_u, const double* x_p, double* b, void* stream);
/* Dot (op 0; :354-396, every DOF of b written, entries of DOF numbers outside iiu and iip set to 0)
 * or Reaction (op 1; :401-433, only the prescribed entries of b written) of a dofval (kind 0) or
 * nodevec (kind 1) */
int gf_matpart_apply(gf_matpart A, int op, int kind, const double* x, double* b, void* stream);
/* MatrixPartitionedSolver::Solve_u  (:545-560, :615-631): x_u = A_uu^-1 (b_u - A_up x_p).  The CG
 * starts from x_u's content when warm != 0, else from zero, and stops when the recursive residual
 * ||r|| <= tol ||b_u - A_up x_p||; *relres receives the true relative residual, *iterations the CG
 * iterations (either may be NULL).  Returns GF_ENOTCONV if the true residual exceeds 10 tol. */
int gf_matpart_solve_u(gf_matpart A, const double* b_u, const double* x_p, double* x_u, int warm, double tol,
                       int64_t maxiter, int64_t* iterations, double* relres, void* stream);
/* MatrixPartitionedSolver::Solve / solve  (:583-613): reads b's unknown and x's prescribed entries,
 * writes x's unknown entries (dofval kind 0, nodevec kind 1); warm: start from x's unknown entries */
int gf_matpart_solve(gf_matpart A, int kind, const double* b, double* x, int warm, double tol, int64_t maxiter,
                     int64_t* iterations, double* relres, void* stream);

/* ================================================================================================
 * Explicit dynamics: one fused velocity-Verlet step of the caller loop
 * docs/examples/dynamics/Elastic-VelocityVerlet/main.cpp:110-157 (asElement -> symGradN_vector ->
 * stress -> int_gradN_dot_tensor2_dV -> assembleNode -> MatrixDiagonal::solve -> updates), on a
 * Quad4 plane-strain mesh with per-element (K, G).  State is dofval FP64 (u, v, a).  Prescribed
 * DOFs carry inv_mass = 0: their acceleration stays 0 and velocity stays as set
 * (MatrixDiagonalPartitioned::solve, MatrixDiagonalPartitioned.h:305-335).
 * ============================================================================================== */

/* coor: device nodevec [nnode, 2] FP64; K_e, G_e: device per-element FP32 moduli (the FP64 path
 * widens them exactly); inv_mass: device dofval FP64; fext: device dofval or NULL (zero). */
int gf_explicit_create(gf_explicit* out, gf_vector v, const double* coor, const float* K_e,
                       const float* G_e, const double* inv_mass, const double* fext, double dt);
int gf_explicit_destroy(gf_explicit x);
/* state arrays (device, [ndof] each); writable between steps */
int gf_explicit_state(gf_explicit x, double** u, double** v, double** a);
/* run nsteps steps; precision GF_F64 (bit-exact with the serial reference order), GF_F32 (FP32
 * strain/stress/force, FP64 assembly and update) or GF_F64_FAST (FP64 element work with FMA).
 * store_fint != 0 keeps the internal force of the last step (one extra [ndof] write). */
int gf_explicit_step(gf_explicit x, int64_t nsteps, int precision, int store_fint, void* stream);
/* Energies of the current state (the per-increment output of Elastic-VelocityVerlet/main.cpp:155-160),
 * enqueued on stream into the device scalars *ekin, *epot: Ekin = 1/2 sum_d mass_d v_d^2 (mass: device
 * dofval, the lumped mass of every DOF, prescribed ones included) and Epot = 1/2 sum_d u_d fint_d, which
 * for the linear elastic law equals sum_q W_q dV_q with W = sigma : eps / 2 (same quadrature).  Needs the
 * last step run with store_fint (GF_ESTATE otherwise).  Deterministic (fixed-order reductions). */
int gf_explicit_energy(gf_explicit x, const double* mass, double* ekin, double* epot, void* stream);
/* internal force (dofval) of the last stored step's displacement, FP64 [ndof] */
int gf_explicit_fint(gf_explicit x, const double** fint);
/* 1 if the structured-grid fused kernel is used (conn is a Regular Quad4 grid), else 0 */
int gf_explicit_is_structured(gf_explicit x, int* structured);
/* which kernels run the step (no counterpart in the reference; a diagnostic): 0 the structured-grid
 * tile kernel, 1 the general mesh fused per node block (every DOF belongs to one node component),
 * 2 the general mesh as element kernel + ordered per-DOF gather (tied DOFs) */
int gf_explicit_path(gf_explicit x, int* path);

/* ================================================================================================
 * Tied-system solver: FP64 Jacobi-preconditioned CG on the operator of MatrixPartitionedTyings
 * (include/GooseFEM/MatrixPartitionedTyings.h:655-674):
 *   K' = A_uu + A_ud C_du + C_du^T A_du + C_du^T A_dd C_du,   K'_up = A_up + A_ud C_dp + ...
 * replacing the direct MatrixPartitionedTyingsSolver (Eigen SimplicialLDLT, :527-675).  A is the
 * stiffness of a linear elastic material (per-element K, G) on the quadrature q, applied
 * matrix-free; v carries conn and the [u, p, d] dofs of Tyings::Periodic (TyingsPeriodic.h:82-121).
 * ============================================================================================== */

/* K_e, G_e: device per-element moduli (FP64).  C_du [nnd x nnu] and C_dp [nnd x nnp] as host CSR
 * (Tyings::Periodic::Cdu / Cdp, TyingsPeriodic.h:272-343; see gf_tyings_matrix). */
int gf_pcg_create(gf_pcg* out, gf_vector v, gf_quad q, const double* K_e, const double* G_e, int64_t nnu,
                  int64_t nnp, const int64_t* cdu_ptr_h, const int64_t* cdu_col_h, const double* cdu_val_h,
                  const int64_t* cdp_ptr_h, const int64_t* cdp_col_h, const double* cdp_val_h);
int gf_pcg_destroy(gf_pcg p);
int gf_pcg_sizes(gf_pcg p, int64_t* nnu, int64_t* nnp, int64_t* nnd);
/* 1 if K' x_u is applied by a fused structured kernel (Quad4 Regular n x n cell, n a multiple of 32,
 * or Hex8 Regular n^3 cell, n a multiple of 8; periodic tyings; identical element geometry), else 0
 * (general element + gather path) */
int gf_pcg_is_structured(gf_pcg p, int* structured);
/* replace the per-element moduli (device, FP64 [nelem] each) and recompute the Jacobi diagonal,
 * enqueued on stream: the tied operator of a new material field without rebuilding the tyings
 * (MatrixPartitionedTyings::assemble + factorize, MatrixPartitionedTyings.h:210-305,655-674) */
int gf_pcg_set_moduli(gf_pcg p, const double* K_e, const double* G_e, void* stream);
/* y_u = K' x_u + K'_up x_p; x_u or x_p may be NULL (zero). */
int gf_pcg_apply(gf_pcg p, const double* x_u, const double* x_p, double* y_u, void* stream);
/* full dofval [x_u, x_p, C_du x_u + C_dp x_p] (the x_d = Cdu X_u + Cdp X_p step of
 * MatrixPartitionedTyingsSolver::solve, MatrixPartitionedTyings.h:547-551) */
int gf_pcg_expand(gf_pcg p, const double* x_u, const double* x_p, double* dofval, void* stream);
/* device pointer to 1 / diag(K') [nnu] (Jacobi preconditioner) */
int gf_pcg_inverse_diagonal(gf_pcg p, const double** invdiag);
/* Solve K' x_u = b_u.  x_u: initial guess in, solution out.  Stops when the recursive residual
 * ||r|| <= tol ||b||; *relres receives the true residual ||b - K' x|| / ||b|| recomputed in FP64,
 * *iterations the CG iterations.  The host checks the device convergence flag every check_every
 * iterations (<= 0: 32).  Returns GF_ENOTCONV if the true residual exceeds 10 tol, GF_EINVAL at
 * once if b_u is not finite.  A start whose residual is not finite or exceeds ||b|| (worse than a
 * zero guess) is replaced by the previous converged solution of this solver, if that one does
 * better than zero, else by zero; gf_pcg_start_used reports which start ran.  Replaces the tied
 * direct solve (MatrixPartitionedTyingsSolver::solve, MatrixPartitionedTyings.h:528-564). */
int gf_pcg_solve(gf_pcg p, const double* b_u, double* x_u, double tol, int64_t maxiter, int64_t check_every,
                 int64_t* iterations, double* relres, void* stream);
#define GF_START_GIVEN 0    /* the caller's x_u */
#define GF_START_PREVIOUS 1 /* the previous converged solution (fallback) */
#define GF_START_ZERO 2     /* zero (fallback) */
/* start used by the last gf_pcg_solve: GF_START_GIVEN, GF_START_PREVIOUS or GF_START_ZERO */
int gf_pcg_start_used(gf_pcg p, int* start);

/* ================================================================================================
 * Learned warm start of the tied PCG (no counterpart in the reference): kernels of a periodic
 * convolutional encoder-decoder, trained and evaluated on the device.  Activations NHWC bfloat16
 * (channel counts multiples of 16); master weights FP32 [9][cout][cin] (tap t = 3 (dy + 1) + dx + 1);
 * packed weights bfloat16.  All reductions are ordered (deterministic).
 * ============================================================================================== */

/* 3x3 circular convolution on the [P = N*H*W] pixel grid.  Input channels [0, c1) from x1 (read
 * at half resolution when up1 != 0: nearest upsampling), [c1, c1 + c2) from x2 (may be NULL if
 * c2 = 0).  wpk: [9][coutp][c1 + c2] (coutp in {32, 64, 128, 256}).  bias: [cout] or NULL.
 * flags: 1 = ReLU, 2 = accumulate into y.  Output channels [0, ysplit) -> y (row stride ysplit),
 * [ysplit, cout) -> y2; y32 (optional) FP32 [P][cout]; mask (optional, [P][cout]) zeroes outputs
 * whose mask value is not > 0.  The data gradient of a convolution is this call with the weights
 * packed tap-flipped and transposed (gf_conv_pack). */
int gf_conv3x3(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int H, int W, const void* wpk,
               int coutp, const float* bias, int cout, int flags, void* y, int ysplit, void* y2, float* y32,
               const void* mask, void* stream);
/* workspace (floats) and split count of gf_conv3x3_wgrad for P pixels, cg gradient channels and
 * cin input channels */
int gf_conv3x3_wgrad_work(int64_t P, int cg, int cin, int* nsplit, int64_t* work_floats);
/* weight gradient dw [9][cout][cin] = sum_p g[p][co] x[p + tap][ci] (same input addressing as
 * gf_conv3x3), bias gradient db [cout] (optional); g has cg channels */
int gf_conv3x3_wgrad(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int H, int W, const void* g,
                     int cg, int cout, int cin, float* work, int64_t work_floats, float* dw, float* db, void* stream);
/* 2x2 average pooling [N, H, W, C] -> [N, H/2, W/2, C] */
int gf_avgpool2(const void* in, int64_t N, int H, int W, int C, void* out, void* stream);
/* fine-grid gradient din[N, H, W, C] (+)= w * dout[N, H/2, W/2, C] (accum != 0 adds), then zeroed
 * where act (optional, fine grid) is not > 0 */
int gf_unpool2(const void* dout, int64_t N, int H, int W, int C, float w, int accum, const void* act, void* din,
               void* stream);
/* coarse gradient of nearest upsampling: out[N, H/2, W/2, C] = sum of each 2x2 block of in[N, H, W, C],
 * zeroed where act (optional, coarse grid) is not > 0 */
int gf_sumpool2(const void* in, int64_t N, int H, int W, int C, const void* act, void* out, void* stream);
/* master weights w [9][cout][cin] -> wf [9][coutp][cin_s] and (optional) wb [9][cinp][cout_s]
 * with wb[t][ci][co] = w[8 - t][co][ci]; zero padded */
int gf_conv_pack(int cout, int cin, const float* w, int coutp, int cin_s, void* wf, int cinp, int cout_s, void* wb,
                 void* stream);
/* Adam (bias-corrected) on n FP32 parameters; step counts from 1 */
int gf_adam(int64_t n, float* w, const float* g, float* m, float* v, float lr, float beta1, float beta2, float eps,
            int64_t step, void* stream);
/* mean squared error of pred vs target ([P][nc] FP32): gradient g [P][cg] bfloat16 (channels >= nc
 * zero, scaled by `scale`), *loss_sum (device) = sum of squared differences; part: >= ceil(P*cg/256)
 * floats of workspace */
int gf_mse(int64_t P, int nc, const float* pred, const float* target, float scale, int cg, void* g, float* part,
           float* loss_sum, void* stream);
/* multi-scale mean squared error of pred vs target on B grids [B][Z][H][W][nc] FP32 (dim 2: Z = 1;
 * dim 3: cells 2^s wide in all three directions): with e = pred - target, mu each grid's and
 * channel's mean of e and m_s the mean of e over the cell of width 2^s (s = 1..S),
 *   L = sum e^2 + lam sum_s (2^dim)^s sum_cells (m_s - mu)^2,
 * which weights error that is smooth over a cell of width 2^s like the plain term and white error
 * (2^dim)^s times less.  Writes g = scale dL/de / (P nc) as bfloat16 into channels 0..nc-1 of
 * [P][cg] (others zero) and *loss_sum (device) = L.  2^S must divide H, W (and Z).  work: at least
 * gf_mse_multiscale_work(...) floats.  No counterpart in the reference (training of the warm start). */
int gf_mse_multiscale_work(int dim, int64_t B, int Z, int H, int W, int nc, int S, int cg, int64_t* nfloats);
int gf_mse_multiscale(int dim, int64_t B, int Z, int H, int W, int nc, int S, float lam, const float* pred,
                      const float* target, float scale, int cg, void* g, float* work, int64_t nwork, float* loss_sum,
                      void* stream);
/* x_u[idx[k]] = gamma * (scale * (grid[k] - grid[k % nc]) + affine[k]) where idx[k] >= 0 */
int gf_grid_to_dofs(int64_t n, int nc, const float* grid, const int32_t* idx, const double* affine, double scale,
                    double gamma, double* xu, void* stream);
/* grid[k] = (x_u[idx[k]] / gamma - affine[k]) / scale (0 where idx[k] < 0); center != 0 subtracts
 * each channel's mean */
int gf_dofs_to_grid(int64_t n, int nc, const double* xu, const int32_t* idx, const double* affine, double scale,
                    double gamma, int center, float* grid, void* stream);
/* 3-D variants on [N, Z, H, W] grids (taps t = 9 (dz + 1) + 3 (dy + 1) + dx + 1, weights [27][..][..],
 * 2x2x2 pooling); arguments as the 2-D calls above */
int gf_conv3x3x3(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W,
                 const void* wpk, int coutp, const float* bias, int cout, int flags, void* y, int ysplit, void* y2,
                 float* y32, const void* mask, void* stream);
int gf_conv3x3x3_wgrad_work(int64_t P, int cg, int cin, int* nsplit, int64_t* work_floats);
int gf_conv3x3x3_wgrad(const void* x1, int c1, int up1, const void* x2, int c2, int64_t P, int Z, int H, int W,
                       const void* g, int cg, int cout, int cin, float* work, int64_t work_floats, float* dw, float* db,
                       void* stream);
int gf_avgpool2x2x2(const void* in, int64_t N, int Z, int H, int W, int C, void* out, void* stream);
int gf_unpool2x2x2(const void* dout, int64_t N, int Z, int H, int W, int C, float w, int accum, const void* act,
                   void* din, void* stream);
int gf_sumpool2x2x2(const void* in, int64_t N, int Z, int H, int W, int C, const void* act, void* out, void* stream);
int gf_conv_pack3(int cout, int cin, const float* w, int coutp, int cin_s, void* wf, int cinp, int cout_s, void* wb,
                  void* stream);
/* network input x [P][cs] bfloat16: channel 0 = scale * (log G - shift), others 0 */
int gf_input_logG(int64_t P, const double* G, double shift, double scale, int cs, void* x, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GOOSEFEM_AMD_H */
