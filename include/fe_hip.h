/*
 * fe_hip.h -- C ABI of libfe_hip.so, the MI355X (gfx950) element-matrix, assembly and solve path.
 *
 * The reference (Dessia-tech/finite_elements v0.2.0) is pure Python with no FFI; its only seams are
 * Python methods.  Each entry point below names the reference function whose work it replaces.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers unless the comment says "host".  The caller owns every
 *     buffer; the library never frees caller memory.  Workspaces use a size query then a caller buffer.
 *   - Return 0 on success; > 0 is a hipError_t passed through; < 0 is a library code (FE_ERR_*).
 *     fe_last_error() returns a thread-local message for the last failure.  No C++ exception crosses
 *     the ABI.
 *   - A handle is bound to one HIP stream and is not re-entrant; separate handles are independent.
 *     Kernels are asynchronous on that stream; only calls returning statistics synchronise.
 *   - Connectivity is SoA: conn[a * n_elem + e] is the node index of local vertex a of element e.
 *   - Element matrices are entry-major: ke[k * n_elem + e].  With sym == 0 k runs over the m*m entries
 *     of the element matrix in row-major order (the order of the reference's ndarray.flatten()); with
 *     sym == 1 over the upper triangle (i <= j) in row-major order.  Local DOF order is
 *     [u1, v1, (w1), u2, v2, ...] as in finite_elements/elements.py:410-412 and 742-747.
 *   - DOF numbering is interleaved: dof = node * dof_per_node + d (finite_elements/core.py:89-108).
 */
#ifndef FE_HIP_H
#define FE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FE_OK 0
#define FE_ERR_ARG (-1)
#define FE_ERR_NOT_CONVERGED (-2)
#define FE_ERR_BREAKDOWN (-3)
#define FE_ERR_UNSUPPORTED (-4)
#define FE_ERR_NOMEM (-5)

/* element kinds */
#define FE_KIND_TRI_ELASTIC 1   /* ElasticityTriangularElement2D, 3 nodes, 2 dof/node */
#define FE_KIND_TET_ELASTIC 2   /* ElasticityTetrahedralElement3D, 4 nodes, 3 dof/node */
#define FE_KIND_TRI_MAGNETIC 3  /* MagneticElement2D, 3 nodes, 1 dof/node */

/* plane modes for FE_KIND_TRI_ELASTIC */
#define FE_PLANE_STRAIN 1
#define FE_PLANE_STRESS 2

/* what to assemble */
#define FE_STIFFNESS 0
#define FE_MASS 1

/* value layouts of an assembled matrix */
#define FE_LAYOUT_CSR 0   /* scalar CSR, columns sorted, the reference's csr_matrix layout */
#define FE_LAYOUT_BSR 1   /* node-block rows: vals[(bptr[i] + q) * dof * dof + di * dof + dj] */
#define FE_LAYOUT_PLANES 2 /* node-block rows, one plane per block entry: vals[(di*dof + dj) * nnzb + q] */

/* preconditioners */
#define FE_PC_JACOBI 1        /* inverse diagonal, one value per DOF */
#define FE_PC_BLOCK_JACOBI 2  /* inverse of each node's dof x dof diagonal block (also the learned one) */

typedef struct fe_handle fe_handle;
typedef struct fe_pattern fe_pattern;

/* Mesh on the device.  coords are SoA; mat0..mat2 are per-element material arrays:
 *   FE_KIND_TRI_ELASTIC : mat0 = Young's modulus, mat1 = Poisson ratio, mat2 = thickness, rho = density
 *   FE_KIND_TET_ELASTIC : mat0 = Young's modulus, mat1 = Poisson ratio,                  rho = density
 *   FE_KIND_TRI_MAGNETIC: mat0 = mu_total (the element's permeability, reluctivity is 1/mu_total)   */
typedef struct {
  int32_t kind;
  int32_t plane_mode;
  int64_t n_elem;
  int64_t n_nodes;
  const double* x;
  const double* y;
  const double* z;
  const int32_t* conn;
  const double* mat0;
  const double* mat1;
  const double* mat2;
  const double* rho;
} fe_mesh;

/* Node graph of a mesh on the device (built by fe_pattern_create, copied to the device by the caller).
 *   bptr/bcol : node adjacency CSR (sorted, self included) = the block pattern of K
 *   inc_ptr/inc : node -> (element, local vertex) incidences, inc = e * npe + a, ascending e per node
 *   bseg_ptr/bseg : for block entry (i, q) the contributing incidences of node i, ascending element,
 *                   bseg = (k_local << 2) | b  with k_local the incidence's rank within node i's list
 *                   and b the local vertex of the column node
 *   blk : node ranges per workgroup of the fused assembly (row-block schedule)
 *   welem_ptr/welem : per workgroup, the ascending distinct elements incident to its node range
 *   inc_sa : per incidence, (slot << 2) | a with slot the element's position in its workgroup's
 *            welem list and a its local vertex (fe_schedule_elements builds these three)          */
typedef struct {
  int64_t n_nodes;
  int64_t nnzb;
  int64_t n_inc;
  int64_t n_blk;
  int32_t max_inc;  /* largest incidence count of one workgroup's node range */
  int32_t threads;  /* workgroup size of the fused assembly: 64, 128, 256, 512 or 1024 (0 = 256) */
  const int32_t* bptr;
  const int32_t* bcol;
  const int32_t* inc_ptr;
  const int32_t* inc;
  const int32_t* bseg_ptr;
  const uint16_t* bseg;
  const int32_t* blk;
  int32_t max_elem; /* largest welem count of one workgroup */
  int32_t producers; /* > 0: the wave-specialised persistent assembly with this many producer waves
                        (needs wg_hdr, welem_conn, max_nbq and max_nseg; 0: not used) */
  const int32_t* welem_ptr;
  const int32_t* welem;
  const uint16_t* inc_sa;
  const int32_t* bdiag; /* per node, the block index of its diagonal entry (-1: node without elements) */
  /* Optional (NULL: derived on the device from the arrays above).  wg_hdr: per workgroup, 12 int32
     {n0, n1, E0, nel, I0, ninc, B0, nbq, c0, nseg, N0, nnode}: its node range [n0, n1), welem list
     start and length, first incidence and count, first block and count, first bseg entry and count,
     and its wnode list's start and length (0, 0 without node lists).
     welem_conn: the node ids of the welem list's elements, [npe][n_welem] (n_welem = welem_ptr[n_blk]). */
  const int32_t* wg_hdr;
  const int32_t* welem_conn;
  int64_t n_welem;
  int32_t max_nodes; /* largest node count of one workgroup's range (0: not given) */
  int32_t max_nbq;   /* largest block count of one workgroup's range (0: not given) */
  int32_t max_nseg;  /* largest bseg entry count of one workgroup's range (0: not given) */
  int32_t max_wnode; /* longest node list of one workgroup (fe_schedule_nodes; 0: not given) */
  /* For the wave-specialised form: the node lists of the element lists (wnode, the workgroup's slice
     in header words 10 and 11: start and length) and per element-list entry the positions of its
     vertices in its workgroup's node list ([n_welem][4]). */
  const int32_t* wnode;
  const uint16_t* welem_local;
  /* For the wave-specialised form: the schedule-resolved block maps of fe_schedule_codes, and one
     map image per range written from them by fe_ws_pack (n_blk images of fe_ws_map_bytes bytes). */
  const uint16_t* wcode;
  const uint16_t* brow;
  const void* wmaps;
} fe_graph;

/* Block sparse matrix (node-block rows) on the device; precision 0 = FP64, 1 = FP32; layout
   FE_LAYOUT_BSR or FE_LAYOUT_PLANES (the solver's layout: every value load is coalesced across the
   lanes that walk a block row).  row_align 0, or 4 for a padded pattern (FE_LAYOUT_PLANES only):
   every block row starts at a multiple of 4 blocks and holds a multiple of 4 blocks, the padding
   blocks having value 0 and the row itself as column, with bcol and vals 16-byte aligned.  The SpMV
   of the solvers then reads 4 (FP32) or 2 (FP64) blocks of a plane with one 16-byte load. */
typedef struct {
  int64_t n_brows;
  int64_t nnzb;
  int32_t dof;
  int32_t precision;
  int32_t layout;
  int32_t row_align;
  const int32_t* bptr;
  const int32_t* bcol;
  const void* vals;
} fe_bsr;

typedef struct {
  double tol;            /* target for the true FP64 relative residual ||b - A x|| / ||b|| */
  int32_t max_iter;
  int32_t precond;       /* FE_PC_* */
  int32_t mixed;         /* 0: FP64 CG; 1: FP32 SpMV/preconditioner with FP64 reliable updates */
  int32_t check_every;   /* iterations per launched chunk between host polls */
  double ru_delta;       /* mixed: reliable update when ||r|| fell by this factor since the last one */
  int32_t profile;       /* 1: after the solve, time 20 back-to-back launches of the iteration's
                            SpMV kernel on the final state with HIP events (stats->spmv_ms) */
  int32_t graph;         /* 1: capture one chunk of iterations in a HIP graph and replay it */
} fe_pcg_opts;

typedef struct {
  int32_t iterations;
  int32_t reliable_updates;
  int32_t status;        /* 0 converged, FE_ERR_NOT_CONVERGED, FE_ERR_BREAKDOWN */
  int32_t restarts;      /* restarts from the true residual after recursive-residual drift */
  double relres;         /* true FP64 relative residual at exit */
  double relres_rec;     /* recursive residual estimate at exit */
  double bnorm;
  double spmv_ms;        /* profile: mean duration of the sampled SpMV kernels */
  int32_t spmv_samples;  /* profile: number of samples */
  int32_t fallbacks;     /* mixed: segments re-solved in FP64 after the FP32 iteration failed */
} fe_pcg_stats;

/* ---- context ---------------------------------------------------------------------------------- */
int fe_version(void);
const char* fe_last_error(void);
int fe_create(fe_handle** h, int device, void* hip_stream);
int fe_destroy(fe_handle* h);
int fe_synchronize(fe_handle* h);

/* ---- element kernels (one thread per element, entry-major output) ------------------------------- */
/* replaces ElasticityTriangularElement2D.elementary_matrix, finite_elements/elements.py:466-511
   (with _b_matrix 395-416 and _d_matrix_plane_strain/_stress 418-453) */
int fe_ke_tri_elastic(fe_handle* h, const double* x, const double* y, const int32_t* conn,
                      const double* young, const double* nu, const double* thickness, int plane_mode,
                      int64_t n_elem, int sym, double* ke);
/* replaces ElasticityTetrahedralElement3D.elementary_matrix, finite_elements/elements.py:809-828
   (with _b_matrix 719-751 and _d_matrix_ 773-796) */
int fe_ke_tet_elastic(fe_handle* h, const double* x, const double* y, const double* z,
                      const int32_t* conn, const double* young, const double* nu, int64_t n_elem,
                      int sym, double* ke);
/* replaces MagneticElement2D.elementary_matrix, finite_elements/elements.py:93-118 */
int fe_ke_tri_magnetic(fe_handle* h, const double* x, const double* y, const int32_t* conn,
                       const double* mu_total, int64_t n_elem, int sym, double* ke);
/* replaces ElasticityTriangularElement2D.elementary_mass_matrix, finite_elements/elements.py:513-536 */
int fe_me_tri(fe_handle* h, const double* x, const double* y, const int32_t* conn, const double* rho,
              const double* thickness, int64_t n_elem, int sym, double* me);
/* replaces ElasticityTetrahedralElement3D.elementary_mass_matrix, finite_elements/elements.py:830-857 */
int fe_me_tet(fe_handle* h, const double* x, const double* y, const double* z, const int32_t* conn,
              const double* rho, int64_t n_elem, int sym, double* me);

/* ---- pattern (host) ------------------------------------------------------------------------------ */
/* replaces the COO index lists of get_row_col_indices / k_matrix_data (finite_elements/analysis.py:
   714-735, 324-339) and the pattern half of scipy's COO->CSR at analysis.py:661.  conn is HOST memory. */
#define FE_PAT_NNZ 0        /* scalar nonzeros of K */
#define FE_PAT_NNZB 1       /* node blocks of K */
#define FE_PAT_NINC 2       /* node-element incidences (= npe * n_elem) */
#define FE_PAT_NCONTRIB 3   /* element-matrix contributions (= m*m * n_elem) */
#define FE_PAT_NBLK 4       /* workgroups of the fused assembly schedule */
#define FE_PAT_MAXINC 5     /* largest incidence count in one workgroup */
#define FE_PAT_NROWS 6      /* n_nodes * dof */
#define FE_PAT_NINFO 8
#define FE_ARR_ROWPTR 0     /* int32 [nrows+1] */
#define FE_ARR_COLIND 1     /* int32 [nnz] */
#define FE_ARR_GATHER 2     /* int32 [ncontrib]: index into entry-major ke of each contribution,
                               grouped by destination CSR entry, ascending element within a group */
#define FE_ARR_SEGPTR 3     /* int32 [nnz+1] */
#define FE_ARR_BPTR 4       /* int32 [n_nodes+1] */
#define FE_ARR_BCOL 5       /* int32 [nnzb] */
#define FE_ARR_INCPTR 6     /* int32 [n_nodes+1] */
#define FE_ARR_INC 7        /* int32 [ninc] */
#define FE_ARR_BSEGPTR 8    /* int32 [nnzb+1] */
#define FE_ARR_BSEG 9       /* uint16 [ninc*npe] */
#define FE_ARR_BLK 10       /* int32 [nblk+1] */
int fe_pattern_create(const int32_t* conn, int npe, int dof, int64_t n_elem, int64_t n_nodes,
                      int sym, int with_gather, fe_pattern** out);
int fe_pattern_schedule(fe_pattern* p, int max_inc_per_block);
int fe_pattern_info(const fe_pattern* p, int64_t* info);
int fe_pattern_copy(const fe_pattern* p, int which, void* host_dst);
int fe_pattern_destroy(fe_pattern* p);
/* Element schedule of the fused assembly for the node ranges blk (n_blk + 1 entries): per workgroup
   the ascending distinct elements incident to its range (welem_ptr: n_blk + 1 entries, welem:
   capacity n_inc) and per incidence its packed (slot << 2 | a) code (inc_sa: n_inc entries);
   max_elem receives the largest element count of one workgroup.  Host arrays; no counterpart in the
   reference (its assembly is a host loop). */
/* Greedy row-block schedule of consecutive nodes (the rule fe_pattern_schedule applies): a range
   takes nodes while it holds at most max_inc incidences and max_nodes nodes.  blk (capacity
   n_nodes + 1) receives the range starts and the end, n_blk the range count, biggest the largest
   range's incidence count.  Host arrays; no counterpart in the reference. */
int fe_schedule_ranges(const int32_t* inc_ptr, int64_t n_nodes, int max_inc, int max_nodes,
                       int32_t* blk, int64_t* n_blk, int32_t* biggest);
int fe_schedule_elements(const int32_t* inc_ptr, const int32_t* inc, int npe, int64_t n_nodes,
                         const int32_t* blk, int64_t n_blk, int32_t* welem_ptr, int32_t* welem,
                         uint16_t* inc_sa, int32_t* max_elem);
/* Node lists of a schedule's element lists (for the wave-specialised assembly): per workgroup w the
   ascending distinct nodes of its elements, wnode[wnode_ptr[w] .. wnode_ptr[w+1]) (wnode capacity
   npe * n_welem), and per element-list entry e the positions of its vertices in that list,
   welem_local[e * 4 + b] (entries b >= npe are 0).  welem_conn is [npe][n_welem]; max_wnode receives
   the longest list.  Host arrays; no counterpart in the reference. */
int fe_schedule_nodes(const int32_t* welem_conn, int npe, int64_t n_welem, const int32_t* welem_ptr,
                      int64_t n_blk, int64_t n_nodes, int32_t* wnode_ptr, int32_t* wnode,
                      uint16_t* welem_local, int32_t* max_wnode);
/* Schedule-resolved block maps (for the wave-specialised assembly): per bseg entry c of block (i, q),
   code[c] = (slot << 4) | (a << 2) | b with slot the element's position in its workgroup's element
   list, a the local vertex of node i and b that of the column node (the bseg and inc_sa codes
   combined, so the block sum reads a record without going through the incidence list); per block
   q, brow[q] = its row node's position in its workgroup's node range, | 0x8000 for the diagonal
   block.  Host arrays (code: bseg's length; brow: nnzb); no counterpart in the reference. */
int fe_schedule_codes(const int32_t* bptr, const int32_t* bcol, const int32_t* bseg_ptr,
                      const uint16_t* bseg, const int32_t* inc_ptr, const uint16_t* inc_sa,
                      int64_t n_nodes, const int32_t* blk, int64_t n_blk, uint16_t* code,
                      uint16_t* brow);

/* ---- assembly ------------------------------------------------------------------------------------ */
/* Two-phase assembly from element matrices: vals[p] = sum over the sorted contribution segment of p.
   LDS-staged segmented reduction, no atomics, bitwise deterministic.  Replaces the value half of
   scipy's COO->CSR sum-duplicates at finite_elements/analysis.py:661 and the dict sum at 387-405. */
int fe_assemble_gather(fe_handle* h, const double* ke, const int32_t* gather, const int32_t* seg_ptr,
                       int64_t nnz, double* vals);
/* Fused assembly: each workgroup owns a node range, computes the element-matrix row blocks of its
   nodes' incidences from coordinates into LDS and reduces them per block entry in ascending element
   order.  No element matrix touches HBM; no atomics; bitwise deterministic.  what = FE_STIFFNESS or
   FE_MASS; layout = FE_LAYOUT_CSR, FE_LAYOUT_BSR or FE_LAYOUT_PLANES.  Replaces k_matrix_data +
   csr_matrix (finite_elements/analysis.py:324-339, 661) and m_matrix_data (357-365). */
int fe_assemble_fused(fe_handle* h, const fe_mesh* mesh, const fe_graph* graph, int what, int layout,
                      double* vals);
/* Map images of the wave-specialised assembly: the size of one range's image for this schedule
   (fe_ws_map_bytes, host only), and the images of all n_blk ranges written into out (device memory,
   zeroed by the caller, n_blk * that size bytes; fe_ws_pack reads the device arrays of g).  No
   counterpart in the reference. */
int fe_ws_map_bytes(const fe_graph* graph, int64_t* bytes);
int fe_ws_pack(fe_handle* h, const fe_graph* graph, void* out);
/* Per-element fields of a solution u (device, interleaved DOFs), component-major out[k * n_elem + e]
   (replaces ElasticityResults.strain / stress / energy_per_element, finite_elements/results.py:689-830,
   and MagneticResults' field, results.py:121-152).  Elasticity: strain and stress in Voigt order (tri
   3: exx, eyy, gxy; tet 6: exx, eyy, ezz, gxy, gyz, gzx) and energy = 1/2 u_e^T Ke u_e; any of the
   three may be NULL.  Magnetic: B = (Bx, By) in strain (2 components), stress and energy NULL. */
int fe_element_fields(fe_handle* h, const fe_mesh* mesh, const double* u, double* strain,
                      double* stress, double* energy);
/* Permute node-chunk values between the CSR and BSR layouts (out-of-place). */
int fe_relayout(fe_handle* h, const int32_t* bptr, int64_t n_nodes, int dof, int to_layout,
                const double* src, double* dst);

/* ---- boundary conditions ------------------------------------------------------------------------- */
/* Dirichlet elimination on a BSR copy of K: rows and columns of fixed DOFs become identity rows.
   Replaces the Lagrange-multiplier augmentation of finite_elements/analysis.py:241-279, 509-543. */
int fe_apply_dirichlet_bsr(fe_handle* h, const fe_bsr* A, const uint8_t* fixed, double* vals);
/* b = f - Kg on free DOFs, b = g on fixed DOFs (Kg computed beforehand with fe_spmv). */
int fe_dirichlet_rhs(fe_handle* h, int64_t n, const uint8_t* fixed, const double* g, const double* f,
                     const double* kg, double* b);

/* ---- sparse kernels -------------------------------------------------------------------------------- */
int fe_spmv(fe_handle* h, const fe_bsr* A, const void* x, void* y);
int fe_spmv_csr(fe_handle* h, int64_t n_rows, const int32_t* rowptr, const int32_t* colind,
                const double* vals, const double* x, double* y);
int fe_cast_f64_to_f32(fe_handle* h, int64_t n, const double* src, float* dst);
/* Copy the values of src into dst->vals converting precision and layout (same pattern). */
int fe_bsr_convert(fe_handle* h, const fe_bsr* src, const fe_bsr* dst);
/* Gather blocks into another pattern of the same rows (e.g. the padded pattern of row_align = 4):
   dst block q = src block src_block[q] converted to dst's precision and layout, or zero where
   src_block[q] < 0.  src_block: device array of dst->nnzb entries.  No counterpart in the reference:
   it builds the solver's iteration copy of the matrix. */
int fe_bsr_repack(fe_handle* h, const fe_bsr* src, const fe_bsr* dst, const int32_t* src_block);
/* Fold src into the pattern of dst (same rows, dof and layout, FP64, unpadded): dst block q =
   sum over k in [seg_ptr[q], seg_ptr[q+1]) of coef[k] * src block src_block[k], in that order.
   With the map of a set of continuity conditions u_s = v u_m (finite_elements_amd/continuity.py) it
   forms T^T K T on the pattern where each slave node's couplings are moved to its master; replaces
   the constraint rows of ContinuityCondition, finite_elements/analysis.py:175-199 (eliminated form). */
int fe_bsr_fold(fe_handle* h, const fe_bsr* src, const fe_bsr* dst, const int32_t* seg_ptr,
                const int32_t* src_block, const double* coef);

/* ---- solver ------------------------------------------------------------------------------------- */
/* Preconditioner setup from the FP64 (Dirichlet-applied) matrix: FE_PC_JACOBI writes n values,
   FE_PC_BLOCK_JACOBI writes n_brows*dof*dof values (the inverse diagonal blocks). */
int fe_precond_setup(fe_handle* h, const fe_bsr* A, int kind, double* out);
/* PCG on the Dirichlet-applied SPD system (replaces spsolve at finite_elements/analysis.py:819-822).
   A64/M64 are FP64; with opts->mixed the FP32 copies A32/M32 run the iteration and FP64 reliable
   updates recompute the true residual.  Without opts->mixed, A32 is NULL (A64 iterates) or an FP64
   iteration copy of A64.  An iteration copy may be on the padded pattern of the same rows
   (row_align = 4, from fe_bsr_repack).  Matrices must be in FE_LAYOUT_PLANES.  x is the initial guess
   on entry.  When the true residual disagrees with the recursive one the iteration restarts from the
   true residual; ten reliable updates without progress (or, in FP32, more than three restarts) end
   the iteration as not converged.  Mixed precision with M64 given: a system whose FP32 iteration
   failed is finished by FP64 PCG from the reached x (plain PCG is the fallback). */
int fe_pcg_workspace(fe_handle* h, const fe_bsr* A, const fe_pcg_opts* opts, int64_t* bytes);
int fe_pcg(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64, const void* M32,
           const double* b, double* x, void* work, int64_t work_bytes, const fe_pcg_opts* opts,
           fe_pcg_stats* stats);

/* Batched PCG over independent segments of one block-diagonal system (replaces one spsolve per design
   variant, finite_elements/analysis.py:819-822): segment g owns node rows [seg_nodes[g],
   seg_nodes[g+1]) (host array, n_seg + 1 entries, 0 .. n_brows) and no block couples two segments.
   Each segment iterates with its own scalars and stops on its own true FP64 residual; seg_stats
   (host, n_seg entries, may be NULL) receives per-segment results, stats the worst status, the
   largest residuals and iteration count. */
typedef struct {
  int32_t iterations;
  int32_t reliable_updates;
  int32_t status;
  int32_t restarts;
  int32_t fallback;      /* 1: the FP32 iteration failed and the segment was finished in FP64 */
  int32_t pad;
  double relres;         /* true FP64 ||b - A x|| / ||b|| of the segment */
  double relres_rec;
  double bnorm;
} fe_pcg_seg_stats;
int fe_pcg_batched_workspace(fe_handle* h, const fe_bsr* A, int32_t n_seg, const int32_t* seg_nodes,
                             const fe_pcg_opts* opts, int64_t* bytes);
int fe_pcg_batched(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                   const void* M32, int32_t n_seg, const int32_t* seg_nodes, const double* b,
                   double* x, void* work, int64_t work_bytes, const fe_pcg_opts* opts,
                   fe_pcg_stats* stats, fe_pcg_seg_stats* seg_stats);

/* ---- two-level preconditioner -------------------------------------------------------------------
   No counterpart in the reference (its solve is a direct spsolve, analysis.py:819-822).  M = M_bj +
   P A_c^-1 P^T: block Jacobi plus a coarse correction over aggregated rigid-body modes.  P has one
   dof x nc block per node (nc = 1, 3, 6 for dof = 1, 2, 3), zero on constrained DOFs, orthonormal per
   aggregate; the pair arrays group the blocks q of the pattern (row pair_row, column bcol[q]) by the
   coarse block pair (agg(row), agg(col)) = pair_ij.  All host-built once per mesh (coarse.py). */
typedef struct {
  int64_t n_nodes;
  int64_t n_agg;
  int64_t n_pairs;
  int32_t dof;
  int32_t nc;
  const int32_t* node_agg;   /* [n_nodes] aggregate of each node */
  const int32_t* agg_ptr;    /* [n_agg + 1] ranges of agg_nodes */
  const int32_t* agg_nodes;  /* nodes grouped by aggregate */
  const int32_t* pair_ptr;   /* [n_pairs + 1] ranges of pair_q / pair_row */
  const int32_t* pair_ij;    /* [n_pairs][2] coarse block pair (I, J) */
  const int32_t* pair_q;     /* block indices grouped by pair (ascending within a pair) */
  const int32_t* pair_row;   /* node row of each pair_q entry */
  const double* P;           /* [n_nodes][dof][nc] */
  const float* P32;          /* FP32 copy for mixed precision (may be NULL otherwise) */
  /* P = Z(x^) T, optional (all NULL: the correction reads P): x^ the node's coordinates centred on
     its aggregate's mean and scaled by its largest extent, Z the rigid-body modes (3D: translations,
     rotations omega x x^ about x, y, z; 2D: translations, rotation; scalar: 1) with the rows of
     constrained DOFs zeroed (fixed_bits: bit e of node i set when DOF i*dof+e is constrained), T the
     aggregate's nc x nc orthonormalising transform.  The correction then evaluates P on the fly. */
  const double* xhat;        /* [n_nodes][3] (2D: z = 0) */
  const float* xhat32;
  const uint8_t* fixed_bits; /* [n_nodes] */
  const double* T;           /* [n_agg][nc][nc] */
  const float* T32;
  /* Block-diagonal coarse space of a batch of independent meshes (n_seg = 0: one system).  Segment s
     (the batched PCG's segment s) owns aggregates [agg_seg[s], agg_seg[s+1]); the coarse operator
     and its inverse are one dense symmetric block per segment, n_s = nc * (its aggregate count)
     square with n_s <= 128, row-major at offset blk_ofs[s]; seg_of_agg[g] is aggregate g's segment.
     Needs xhat, T and fixed_bits (the correction evaluates P from them). */
  int32_t n_seg;
  int32_t max_block;         /* largest n_s */
  int64_t blk_total;         /* blk_ofs[n_seg] */
  const int32_t* agg_seg;    /* [n_seg + 1] */
  const int32_t* seg_of_agg; /* [n_agg] */
  const int64_t* blk_ofs;    /* [n_seg + 1] */
} fe_coarse;
/* Leading dimension of the dense coarse matrices: n_agg * nc rounded up to a multiple of 64. */
int64_t fe_coarse_ld(int64_t n_coarse);
/* A_c = P^T A P (FP64 planes A) into Ac [ld][ld] (identity in the padding); with invert, Ac is
   replaced by A_c^-1 (blocked Gauss-Jordan, FP64, symmetrised) and Ac32 (may be NULL) receives its
   FP32 copy; scratch holds ld * ld + 2 * 64 * 64 doubles.  Block-diagonal space (n_seg > 0): Ac and
   Ac32 hold the per-segment blocks (blk_ofs[n_seg] values), each inverted by one workgroup in LDS;
   scratch is not used. */
int fe_coarse_operator(fe_handle* h, const fe_bsr* A, const fe_coarse* C, int invert, double* Ac,
                       float* Ac32, double* scratch);
/* u += P Ainv P^T r in FP64 (precision 0) or FP32 (1); work: 2 * round_up(n_coarse * elem, 256) bytes.
   A block-diagonal space takes its FP64 blocks as Ainv at either precision. */
int fe_coarse_apply(fe_handle* h, const fe_coarse* C, int precision, const void* Ainv, const void* r,
                    void* u, void* work);
/* fe_coarse_apply without a workspace, for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
   reference.  The coarse vectors live in a static device array of the library (1 MiB per GPU, held
   until the process ends), so n_agg * nc must be <= FE_COARSE_NOALLOC_MAX (FE_ERR_ARG otherwise);
   the call allocates nothing.  Every call on one GPU shares that array: calls that overlap on one GPU
   (several streams or threads) race and must be serialised by the caller. */
#define FE_COARSE_NOALLOC_MAX 65536
int fe_coarse_apply_noalloc(fe_handle* h, const fe_coarse* C, int precision, const void* Ainv,
                            const void* r, void* u);
/* PCG with the two-level preconditioner (one system): as fe_pcg, with the coarse correction added to
   the block-Jacobi (or learned) M after every update of the residual.  Ainv64/Ainv32 from
   fe_coarse_operator (Ainv32 for opts->mixed).  fe_pcg_batched_coarse: the same for the segments of
   fe_pcg_batched, with a block-diagonal coarse space whose segments are the PCG's; its coarse blocks
   are applied in FP64 at every precision (Ainv64 required, Ainv32 unused) (no counterpart in the
   reference, whose plates are solved one by one by spsolve). */
int fe_pcg_coarse_workspace(fe_handle* h, const fe_bsr* A, const fe_coarse* C, const fe_pcg_opts* opts,
                            int64_t* bytes);
int fe_pcg_coarse(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64, const void* M32,
                  const fe_coarse* C, const void* Ainv64, const void* Ainv32, const double* b, double* x,
                  void* work, int64_t work_bytes, const fe_pcg_opts* opts, fe_pcg_stats* stats);
int fe_pcg_batched_coarse_workspace(fe_handle* h, const fe_bsr* A, int32_t n_seg, const int32_t* seg_nodes,
                                    const fe_coarse* C, const fe_pcg_opts* opts, int64_t* bytes);
int fe_pcg_batched_coarse(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                          const void* M32, int32_t n_seg, const int32_t* seg_nodes, const fe_coarse* C,
                          const void* Ainv64, const void* Ainv32, const double* b, double* x, void* work,
                          int64_t work_bytes, const fe_pcg_opts* opts, fe_pcg_stats* stats,
                          fe_pcg_seg_stats* seg_stats);

/* ---- eigen-solve ---------------------------------------------------------------------------------
   Batched LOBPCG for K x = lambda M x (replaces the eigsh / dense-inverse eigs calls of
   finite_elements/analysis.py:741-796).  Variants share the pattern of K; variant v uses the values
   at K->vals + v * k_stride (M likewise; stride 0 = shared) and the block-Jacobi blocks
   T + v * t_stride ([n_brows][dof][dof], FP64; T = NULL: no preconditioner, the choice for
   largest = 1).  Matrices FP64, K and M in the same layout: FE_LAYOUT_BSR (faster: a row's blocks are
   read contiguously) or FE_LAYOUT_PLANES.  X: [n_variants][n][block] row-major, the initial
   block on entry (zero on fixed DOFs), the Ritz vectors on return (M-orthonormal, ascending eigenvalue
   for largest = 0).  lambda/resid: [n_variants][block]; resid = ||K x - l M x|| / (|l| ||M x||).
   A variant stops when its n_wanted first pairs reach tol. fixed (n bytes, may be NULL) zeroes the
   preconditioned residual on constrained DOFs, which keeps the iteration in the free subspace when K
   and M carry identity / zero rows there. */
typedef struct {
  int32_t n_variants;
  int32_t block;        /* columns of the iterated block: multiple of 4, 4..32 */
  int32_t n_wanted;     /* <= block */
  int32_t largest;      /* 0: smallest eigenvalues, 1: largest */
  int32_t max_iter;
  int32_t check_every;  /* iterations per launched chunk between host polls */
  double tol;
  int64_t k_stride;     /* values between consecutive variants of K (0: one K for all) */
  int64_t m_stride;
  int64_t t_stride;     /* doubles between consecutive variants of T */
} fe_lobpcg_opts;

typedef struct {
  int32_t max_iterations;
  int32_t unconverged;  /* variants that reached max_iter */
  int32_t failed;       /* variants whose basis lost rank below the block size */
  int32_t status;       /* FE_OK, FE_ERR_NOT_CONVERGED or FE_ERR_BREAKDOWN */
} fe_lobpcg_stats;

int fe_lobpcg_workspace(fe_handle* h, const fe_bsr* K, const fe_lobpcg_opts* opts, int64_t* bytes);
int fe_lobpcg(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T, const uint8_t* fixed,
              double* X, double* lambda, double* resid, int32_t* iters, int32_t* variant_status,
              void* work, int64_t work_bytes, const fe_lobpcg_opts* opts, fe_lobpcg_stats* stats);
/* fe_lobpcg with the two-level preconditioner T + P A_c^-1 P^T (no counterpart in the reference, whose
   eigsh factorises K): C is the coarse space of the constrained mesh (shared by every variant), Ainv
   the per-variant FP64 coarse inverses [ld][ld] from fe_coarse_operator, ainv_stride doubles apart. */
int fe_lobpcg_coarse_workspace(fe_handle* h, const fe_bsr* K, const fe_coarse* C,
                               const fe_lobpcg_opts* opts, int64_t* bytes);
int fe_lobpcg_coarse(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T,
                     const uint8_t* fixed, const fe_coarse* C, const double* Ainv, int64_t ainv_stride,
                     double* X, double* lambda, double* resid, int32_t* iters, int32_t* variant_status,
                     void* work, int64_t work_bytes, const fe_lobpcg_opts* opts, fe_lobpcg_stats* stats);
/* Batched symmetric eigen-decomposition (parallel cyclic Jacobi, one workgroup per matrix, FP64):
   A [n_mat][r][r] (symmetrised on read), w [n_mat][r] eigenvalues (unordered), V [n_mat][r][r]
   eigenvectors in columns; r <= 96.  The Rayleigh-Ritz kernel of fe_lobpcg uses the same routine. */
int fe_sym_eig(fe_handle* h, int n_mat, int r, const double* A, double* w, double* V);

/* ---- learned preconditioner ----------------------------------------------------------------------
   No counterpart in the reference (its solve is a direct spsolve): a per-node graph network over the
   node graph of K whose output scales each block-Jacobi block, M_i = diag(s_i) K_ii^-1 diag(s_i),
   s_i = exp(clamp(o_i, -4, 4)).  3 message-passing layers of width 32 on 8 scale-free node features.
   Parameters are one FP32 vector of fe_gnn_num_params(dof) entries.
   acts: [3][n_nodes][32] layer outputs; aggs: [n_nodes][8 + 32 + 32] neighbour means (for backward).
   fe_gnn_loss_grad accumulates, for one residual/error pair (r, e) with A e = r and z = M r + zc, the
   loss into *loss_acc and dL/do into dlo: kind 0, L = 1 - (r.z)^2 / ((z.Az)(r.e)) (energy left after
   an optimally damped step); kind 1, L = |e - z|_A^2 / |e|_A^2 (the undamped step).  zc (may be NULL) is a part
   of the preconditioned residual that does not depend on the network: the coarse correction
   P A_c^-1 P^T r when the learned blocks are the smoother of the two-level preconditioner. */
int fe_gnn_num_params(int dof); /* returns the count (>= 0) or FE_ERR_ARG */
int fe_gnn_features(fe_handle* h, const fe_bsr* A, const uint8_t* fixed, float* feat, double* scratch);
int fe_gnn_forward(fe_handle* h, int64_t n_nodes, int dof, const int32_t* bptr, const int32_t* bcol,
                   const float* params, const float* feat, float* acts, float* aggs,
                   const double* binv, float* o, double* M);
int fe_gnn_loss_workspace(int64_t n_dofs, int64_t* doubles);
int fe_gnn_loss_grad(fe_handle* h, const fe_bsr* A, const double* binv, const double* M,
                     const float* o, const double* r, const double* e, const double* zc, double* work,
                     double* loss_acc, float* dlo, int kind);
int fe_gnn_backward_workspace(int64_t n_nodes, int64_t* floats);
int fe_gnn_backward(fe_handle* h, int64_t n_nodes, int dof, const int32_t* bptr, const int32_t* bcol,
                    const float* params, const float* feat, const float* acts, const float* aggs,
                    const float* dlo, float* work, float* grad);
int fe_gnn_adam(fe_handle* h, int n, float* params, const float* grad, float* m, float* v, float lr,
                float beta1, float beta2, float eps, int step, float grad_scale);

#ifdef __cplusplus
}
#endif

#endif /* FE_HIP_H */
