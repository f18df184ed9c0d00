/*
 * mife.h - C ABI of the MI355X-native finite-element assembly and sparse-solve
 * library (libmi_fe.so) that stands in for OpenCMISS-Iron's
 *   FiniteElementCalculate -> EquationsMatrices element-add -> DistributedMatrix CSR
 *   -> Dirichlet reduction -> KSP (CG / GMRES + Jacobi) -> back-substitution
 * path.  Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status:
 *   0  MIFE_OK                         (CMFE_NO_ERROR, bindings/generate_bindings/c.py:31)
 *  -1  MIFE_ERR_POINTER_IS_NULL        (CMFE_POINTER_IS_NULL, c.py:32)
 *  -3  MIFE_ERR_COULD_NOT_ALLOCATE     (CMFE_COULD_NOT_ALLOCATE_POINTER, c.py:34)
 *  >0  errors raised by the library itself, as iron's FlagError (err=1,
 *      src/base_routines.F90:673-741) with a message from mife_last_error().
 *
 * Buffers are device pointers owned by the caller unless stated otherwise;
 * work is ordered on the caller's hipStream_t (passed as void*; NULL = default
 * stream).  Index arrays are 0-based int32 (iron's CSR is 1-based; the
 * conversion is a constant -1).  Calls that read results back to the host
 * (iteration counts, norms) say so.  The vector entry points (gather, solver
 * RHS, field update, back-substitution) return MIFE_OK at once for a zero
 * count, and their array arguments may then be NULL.
 */
#ifndef MIFE_H
#define MIFE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIFE_OK 0
#define MIFE_ERR_POINTER_IS_NULL (-1)
#define MIFE_ERR_COULD_NOT_ALLOCATE (-3)
#define MIFE_ERR 1
#define MIFE_ERR_ARG 2
#define MIFE_ERR_HIP 3
#define MIFE_ERR_NOT_CONVERGED 4
#define MIFE_ERR_PATTERN 5

/* Value precision selector for CSR values and element kernels. */
#define MIFE_F64 0
#define MIFE_F32 1

/* Solver-matrix storage formats. */
#define MIFE_FMT_CSR 0
#define MIFE_FMT_SELL 1
#define MIFE_FMT_SELL3 3

/* A square solver matrix (values passed separately, FP64 or FP32).
 * CSR: ptr = row_ptr[n+1], col = column indices, perm = NULL.
 * SELL (slices of 64 rows, one row per lane of a wavefront): slice row k is
 * original row perm[k]; ptr = slice_ptr[n_slices+1] (entry offsets, multiples
 * of 64); entry j of row k of slice s sits at ptr[s] + 64 j + (k - 64 s); col
 * holds column indices in the slice numbering (padding: value 0).  Solvers take
 * b and x in the original numbering and permute internally.
 * Compressed SELL: when col_base (per-slice base column, [n_slices]) is not
 * NULL, col points to uint16_t offsets and the column of an entry of slice s is
 * col_base[s] + offset (mife_sell_compress); CSR ignores col_base.
 * SELL schedule: when slice_order is not NULL it holds 8 ranges of order_span
 * slots (mife_sell_order): range k lists the slices of the k-th eighth of the
 * matrix, longest first (16-byte aligned).  A slot is four int32: the slice, its first entry
 * (ptr[slice]), its entries per row and its base column (col_base[slice], 0
 * without compression); padding slots are {-1, 0, 0, 0}.  SELL products then
 * give range k to the blocks that the dispatcher places on XCD k, start the
 * longest slices first and read a slice's bounds in one load; NULL keeps the
 * plain slice order.  CSR ignores both fields.
 * SELL3 (three unknowns per node, the three rows of a node alike): the SELL
 * layout over node rows with 3x3 blocks as entries.  n = 3 x node rows; vector
 * position 3k + c (node row k in slice order, component c) is original row
 * perm[3k + c]; ptr = slice_ptr[n_slices+1] in blocks (multiples
 * of 64), n_slices = ceil(n / 192); block j of slice node row k of slice s
 * (b = ptr[s] + 64 j + (k - 64 s)) has node column col_base[s] + col[b]
 * (uint16_t, col_base required) and value t = 3 r + c (row r, column
 * component c) at values[9 (ptr[s] + 64 j) + 64 t + (k - 64 s)]; padding
 * blocks hold zeros.  nnz = stored values (9 per block).  slice_order as
 * SELL, with the base and length in blocks. */
typedef struct mife_matrix {
  int n;
  int format;
  const int32_t* ptr;
  const int32_t* col;
  const int32_t* perm;
  int n_slices;
  int64_t nnz; /* stored entries (SELL: ptr[n_slices], padding included); sizes launches only */
  const int32_t* col_base;
  const int32_t* slice_order;
  int order_span;
} mife_matrix;

/* ---------------------------------------------------------------- runtime */

/* Library version string (static storage). */
const char* mife_version(void);

/* Copies the calling thread's last error message into buf (NUL-terminated).
 * Replaces cmfe_ExtractErrorMessage (bindings/python/opencmiss/iron/_utils.in.py:66-77). */
int mife_last_error(char* buf, int n);

/* Selects the HIP device and warms the runtime.  Replaces the device-free
 * cmfe_Initialise_ (src/cmiss.F90:237). */
int mife_init(int device);

/* ------------------------------------------------- sparsity and DOF maps */

/* Builds the FEM CSR sparsity of an equations matrix on the GPU:
 * row d holds the sorted unique union of the DOF lists of every element whose
 * DOF list contains d.  Replaces EquationsMatrix_StructureCalculate (FEM/CSR,
 * src/equations_matrices_routines.F90:4399-4461) with LIST_REMOVE_DUPLICATES
 * (src/lists.F90:2652-2700).
 *   elem_dofs [n_elem*n_edof]  element DOF lists (device)
 *   row_ptr   [n_dofs+1]       output (device)
 * Pass 1 (col == NULL): fills row_ptr and writes the nnz to *nnz_out (host read).
 * Pass 2 (col != NULL): fills col[nnz] (device). work must hold
 * mife_csr_pattern_work_bytes(n_elem, n_edof, n_dofs) bytes and be preserved
 * between the two passes. */
size_t mife_csr_pattern_work_bytes(int n_elem, int n_edof, int n_dofs);
int mife_csr_pattern(const int32_t* elem_dofs, int n_elem, int n_edof, int n_dofs,
                     int32_t* row_ptr, int32_t* col, int64_t* nnz_out,
                     void* work, void* stream);

/* Per-element CSR position map: pos[e][i][j] = position of column
 * elem_dofs[e][j] in row elem_dofs[e][i].  Replaces the per-entry
 * MATRIX_STORAGE_LOCATION_FIND bisection (src/matrix_vector.F90:1724-1854)
 * done inside every MATRIX_VALUES_ADD_DP2 (:3119) with a one-off table.
 * Returns MIFE_ERR_PATTERN if an entry is missing (checked on device, read back). */
int mife_scatter_map(const int32_t* elem_dofs, int n_elem, int n_edof,
                     const int32_t* row_ptr, const int32_t* col,
                     int32_t* pos, void* stream);

/* Greedy element colouring on the host (no two elements of one colour share a
 * DOF), so that coloured scatter-adds are race-free and deterministic.
 * elem_dofs is a HOST array here.  colour[n_elem] is written; returns the
 * number of colours in *n_colours. */
int mife_colour_elements(const int32_t* elem_dofs_host, int n_elem, int n_edof,
                         int n_dofs, int32_t* colour_host, int* n_colours);

/* Reduced (solver) pattern over the free DOFs: keeps entries whose row and
 * column are both free, renumbers them by rank among free DOFs and records the
 * source position of each kept entry.  Replaces SOLVER_MAPPING_CALCULATE row and
 * column inclusion (src/solver_mapping_routines.F90:422-547, :1445-1511) and
 * SOLVER_MATRIX_STRUCTURE_CALCULATE (src/solver_matrices_routines.F90:1612-2431).
 *   new_idx [n]    rank of each DOF among free DOFs, -1 if fixed (device)
 * Pass 1 (s_col == NULL): s_row_ptr[n_free+1] and *nnz_out (host read).
 * Pass 2: s_col[nnz_f], src_pos[nnz_f]. */
int mife_reduce_pattern(int n, const int32_t* row_ptr, const int32_t* col,
                        const int32_t* new_idx, int n_free,
                        int32_t* s_row_ptr, int32_t* s_col, int32_t* src_pos,
                        int64_t* nnz_out, void* work, void* stream);

/* Row permutation of a SELL-64 matrix: positions are cut into windows of
 * sigma (1..1024) rows; inside a window the rows go by decreasing length, ties
 * by position.  row_at[p] is the row at position p (NULL: row p); perm[k] is
 * the row stored at slice position k.  No reference counterpart (the
 * reference's solver matrix is a PETSc AIJ matrix). */
int mife_sell_permutation(int n, const int32_t* row_ptr, const int32_t* row_at, int sigma, int32_t* perm,
                          void* stream);

/* SELL-64 layout of a CSR matrix: given the row permutation (perm, iperm) and
 * slice_ptr computed from the row lengths, writes sell_col (slice numbering)
 * and sell_src (position of each stored entry in the CSR values, -1 for
 * padding) so that mife_gather_values fills SELL values straight from the
 * equations matrix.  No reference counterpart: the reference's solver matrix
 * is a PETSc AIJ (CSR) matrix (src/distributed_matrix_vector.F90:2465-2581). */
int mife_sell_build(int n, const int32_t* row_ptr, const int32_t* col, const int32_t* src_pos,
                    const int32_t* perm, const int32_t* iperm, int n_slices,
                    const int32_t* slice_ptr, int32_t* sell_col, int32_t* sell_src,
                    void* stream);

/* 16-bit SELL columns: col_base[s] = smallest column of slice s, col16[e] =
 * sell_col[e] - col_base[s].  *ok (host) is 1 when every slice spans fewer than
 * 65,536 columns (then col16/col_base are valid), else 0.  Synchronises the
 * stream.  Halves the column bytes the SELL products stream. */
int mife_sell_compress(int n_slices, const int32_t* slice_ptr, const int32_t* sell_col,
                       int32_t* col_base, uint16_t* col16, int* ok, void* stream);

/* Length (int32 entries) of the slice_order array of a SELL matrix with
 * n_slices slices: 4 * 8 * order_span, order_span = 4 * ceil(ceil(n_slices / 8) / 4). */
int mife_sell_order_len(int n_slices);

/* SELL schedule (see mife_matrix): splits the slices into 8 contiguous ranges
 * and lists each range by decreasing slice length (ties by slice index) as
 * {slice, first entry, entries per row, base column} slots, padding with
 * {-1, 0, 0, 0}.  col_base: the compressed storage's per-slice base columns,
 * or NULL (base 0).  order has mife_sell_order_len(n_slices) int32 entries
 * (device); *order_span (host) receives the per-range slot count.  The ranking
 * runs on the host from copies of slice_ptr and col_base; synchronises the
 * stream.  No reference counterpart (PETSc's MatMult walks CSR rows in order). */
int mife_sell_order(int n_slices, const int32_t* slice_ptr, const int32_t* col_base, int32_t* order,
                    int* order_span, void* stream);

/* ------------------------------------------------------------- assembly */

/* Element-kernel geometry and basis description shared by the assembly calls.
 * Tables are device arrays in Gauss-point-major order. */
typedef struct mife_element_desc {
  int dim;          /* number of xi = number of x (2 or 3) */
  int n_nodes;      /* element nodes of the (single) basis: 4, 8, 9, 27 */
  int n_gauss;      /* Gauss points */
  const double* dN; /* [n_gauss][n_nodes][dim] dphi/dxi (GAUSS_BASIS_FNS(:,PART_DERIV_Si,:)) */
  const double* N;  /* [n_gauss][n_nodes] phi (GAUSS_BASIS_FNS(:,NO_PART_DERIV,:)) or NULL */
  const double* w;  /* [n_gauss] Gauss weights */
} mife_element_desc;

/* Scalar Laplace / generalised-Laplace / constant-source Poisson assembly with
 * deterministic coloured scatter into CSR values.  Replaces, per element,
 * LaplaceEquation_FiniteElementCalculate (src/Laplace_equations_routines.F90:482-747),
 * POISSON_EQUATION_FINITE_ELEMENT_CALCULATE constant source
 * (src/Poisson_equations_routines.F90:3347-3418) and EquationsMatrices_ElementAdd
 * (src/equations_matrices_routines.F90:1140-1265).
 *   coords   [n_nodes_geom][dim] FP64 node coordinates (device)
 *   elem_nodes [n_elem][n_nodes] geometric node of each element node (device)
 *   elem_dofs  [n_elem][n_nodes] row/column DOF of each element node (device)
 *   kappa    [n_elem] element conductivity or NULL (=1)
 *   source   [n_elem] element source f or NULL (no source vector)
 *   pos      [n_elem][n_nodes][n_nodes] scatter map
 *   order    [n_elem] element ids grouped by colour; colour_ptr[n_colours+1]
 *            (host array) delimits each colour; order == NULL means atomics.
 *   vals     CSR values (FP64 or FP32 per `prec`), accumulated into (+=)
 *   src_vec  FP64 source vector, accumulated into (+=), may be NULL */
int mife_assemble_scalar(const mife_element_desc* desc, int n_elem,
                         const double* coords, const int32_t* elem_nodes,
                         const int32_t* elem_dofs, const double* kappa,
                         const double* source, const int32_t* pos,
                         const int32_t* order, const int32_t* colour_ptr_host,
                         int n_colours, int prec, void* vals, double* src_vec,
                         void* stream);

/* 3-D linear elasticity (orthotropic C from E1,E2,E3,v13,v23,v12 per element),
 * element DOFs component-blocked [u1(1..nn),u2,u3].  Replaces
 * LINEAR_ELASTICITY_FINITE_ELEMENT_CALCULATE (src/linear_elasticity_routines.F90:1109-1403)
 * and LINEAR_ELASTICITY_TENSOR (:1410-1510) plus EquationsMatrices_ElementAdd.
 *   material [n_elem][6] FP64 (E1,E2,E3,v13,v23,v12)
 *   elem_dofs [n_elem][3*n_nodes] */
int mife_assemble_elastic(const mife_element_desc* desc, int n_elem,
                          const double* coords, const int32_t* elem_nodes,
                          const int32_t* elem_dofs, const double* material,
                          const int32_t* pos, const int32_t* order,
                          const int32_t* colour_ptr_host, int n_colours, int prec,
                          void* vals, void* stream);

/* Structured assembly on a generated regular 3-D mesh: n_elem[3] elements per
 * direction, elements and nodes numbered x fastest (iron's generated mesh,
 * src/generated_mesh_routines.F90:1568-1612), one component (a scalar field:
 * Laplace / generalised Laplace / Poisson, coef = per-element conductivity or
 * NULL, source = per-element f or NULL) or three (linear elasticity, coef =
 * material [E][6], no source) in separated DOF order, trilinear (8 nodes, 2x2x2
 * Gauss) or, for one component, triquadratic (27 nodes, 3x3x3 Gauss) elements.
 * Replaces the element loop with EquationsMatrices_ElementAdd
 * (src/equations_matrices_routines.F90:1140-1265, MATRIX_VALUES_ADD_DP2
 * src/matrix_vector.F90:3119): the element kernels store every element matrix in
 * `stage` (one launch), then each CSR row (row_ptr over n_rows = n_comp x nodes
 * rows, iron's sorted pattern from mife_csr_pattern on the generated mesh's
 * element DOFs -- the full grid stencil; the caller guarantees it, as the row
 * structure is derived from the grid indices) is written whole, one x-line of
 * rows per wavefront: entry by entry the element entries that meet there, summed
 * in the element-colour order
 * of the coloured scatter, so vals (and src_vec) equal bit for bit what
 * mife_assemble_scalar / mife_assemble_elastic with the structured colouring add
 * into zeroed arrays.  vals and src_vec are overwritten, not added to; no
 * position map is read.  stage: mife_assemble_grid_stage_bytes(E, nodes per
 * element, n_comp, prec) bytes of device memory (the FP64 size serves both). */
size_t mife_assemble_grid_stage_bytes(int n_elem, int n_nodes, int n_comp, int prec);
int mife_assemble_grid(const mife_element_desc* desc, const int32_t* n_elem, int n_comp,
                       const double* coords, const int32_t* elem_nodes, const double* coef,
                       const double* source, const int32_t* row_ptr, int64_t n_rows, int prec,
                       void* vals, double* src_vec, void* stage, size_t stage_bytes, void* stream);

/* Device structures of a generated box (a served request's mesh): node
 * coordinates, element connectivity, the equations CSR, the Dirichlet
 * reduction and the SELL-64 storage of the reduced matrix, written from closed
 * forms of the grid indices.  Replaces, for a box, the generated regular mesh
 * and DOF numbering (src/generated_mesh_routines.F90:1568-1612, :3217-3384;
 * src/field_routines.F90:10056), EquationsMatrix_StructureCalculate
 * (src/equations_matrices_routines.F90:4316-4602) and the solver mapping and
 * matrix structure (src/solver_mapping_routines.F90:422-547, :1445-1511;
 * src/solver_matrices_routines.F90:1612-2431).  Numbering as the generic path:
 * nodes and elements x fastest, DOFs in separated order (c N + node), CSR rows
 * sorted; solver rows and columns are the free DOFs ascending.  Dirichlet
 * planes x = 0 (fix_lo, value u_lo) and x = x_max (fix_hi, value u_hi), every
 * component; the box is [0, extent].
 * mife_box_sizes fills the counts on the host (no device work); mife_box_build
 * writes the arrays on `stream` without allocating or synchronising:
 *   coords [n_nodes][3] (may be NULL), elem_nodes [n_elem][nodes_per_elem],
 *   row_ptr [n_dofs + 1], col [nnz], free_idx [n_free], new_idx [n_dofs] (-1:
 *   fixed), u0 [n_dofs] (Dirichlet values, 0 on free DOFs), slice_ptr
 *   [n_slices + 1], sell_src [stored] (equations-CSR position, -1: padding) and
 *   either sell_col16 [stored] + col_base [n_slices] (col16 = 1) or sell_col
 *   [stored].  Every slice has width `width` (the longest reduced row); padding
 *   entries hold the row's own column and no source.
 * With block = 1 and n_comp = 3 (counts.block = 1) the free DOFs are numbered
 * node by node (solver index 3 k + c for the k-th free node, x fastest) and
 * the reduced matrix is SELL3 (MIFE_FMT_SELL3): n_slices = ceil(free nodes /
 * 64), width in node blocks, slice_ptr in blocks, sell_col16 [stored / 9] node
 * columns with col_base [n_slices], sell_src [stored] (per block value). */
typedef struct {
  int32_t n_elem[3];
  int32_t order;   /* 1 trilinear, 2 triquadratic */
  int32_t n_comp;  /* 1 (scalar field) or 3 (displacement) */
  int32_t fix_lo, fix_hi;
  int32_t block;   /* 1: SELL3 storage for n_comp = 3 (when its 16-bit columns fit) */
  double extent[3];
  double u_lo, u_hi;
} mife_box_desc;

typedef struct {
  int64_t n_nodes, n_elem, nodes_per_elem, n_dofs, nnz;
  int64_t n_free, red_nnz, n_slices, width, stored;
  int32_t col16, block; /* block: the reduced matrix is SELL3 */
} mife_box_counts;

typedef struct {
  double* coords;
  int32_t* elem_nodes;
  int32_t* row_ptr;
  int32_t* col;
  int32_t* free_idx;
  int32_t* new_idx;
  double* u0;
  int32_t* slice_ptr;
  int32_t* sell_col;
  uint16_t* sell_col16;
  int32_t* col_base;
  int32_t* sell_src;
} mife_box_arrays;

int mife_box_sizes(const mife_box_desc* desc, mife_box_counts* out);
int mife_box_build(const mife_box_desc* desc, const mife_box_arrays* arrays, void* stream);

/* Batched assembly: the CSR values of many independent trilinear meshes (the
 * requests a server admits together) in one launch per colour and element
 * kind.  Each job is one mesh with the arguments of mife_assemble_scalar
 * (kind MIFE_JOB_SCALAR: coef = kappa or NULL, optional source) or
 * mife_assemble_elastic (MIFE_JOB_ELASTIC: coef = material); its elements are
 * taken in its colour order, so its values equal those of its own
 * mife_assemble_* call bit for bit.  Jobs must not share output arrays.
 * Trilinear elements with 2x2x2 Gauss points only (scalar_desc / elastic_desc:
 * dim 3, 8 nodes, 8 points; either may be NULL when no job of its kind is
 * given).  The host table (table_host, pageable or pinned) and the device
 * table (work) hold mife_assemble_batch_work_bytes(n_jobs, max_colours) bytes;
 * the call synchronises the stream once after copying the table.  No
 * reference counterpart: iron assembles one equations set per call of
 * EquationsSet_AssembleStaticLinearFEM (src/equations_set_routines.F90:1203-1364). */
#define MIFE_JOB_SCALAR 0
#define MIFE_JOB_ELASTIC 1

typedef struct mife_assembly_job {
  int32_t kind;
  int32_t n_elem;
  const double* coords;           /* [nodes][3] */
  const int32_t* elem_nodes;      /* [n_elem][8] */
  const int32_t* elem_dofs;       /* scalar with a source: [n_elem][8] */
  const double* coef;             /* scalar: kappa [n_elem] or NULL; elastic: material [n_elem][6] */
  const double* source;           /* scalar: f [n_elem] or NULL */
  const int32_t* pos;             /* scatter map [n_elem][nd][nd] */
  const int32_t* order;           /* element ids grouped by colour */
  const int32_t* colour_ptr_host; /* [n_colours + 1], host memory */
  int32_t n_colours;
  int32_t pad;
  void* vals;                     /* CSR values (FP64 or FP32 per prec), accumulated into */
  double* src_vec;                /* scalar: FP64 source vector or NULL */
} mife_assembly_job;

size_t mife_assemble_batch_work_bytes(int n_jobs, int max_colours);
int mife_assemble_batch(const mife_element_desc* scalar_desc, const mife_element_desc* elastic_desc,
                        int n_jobs, const mife_assembly_job* jobs, int max_colours, int prec,
                        void* table_host, void* work, void* stream);

/* Mooney-Rivlin finite elasticity (3-D, FP64). */
#define MIFE_MR_DEVIATORIC 0 /* EQUATIONS_SET_MOONEY_RIVLIN_SUBTYPE branch */
#define MIFE_MR_CAUCHY 1     /* EQUATIONS_SET_NO_SUBTYPE general branch */
#define MIFE_MR_NEARLY_INCOMPRESSIBLE 2 /* EQUATIONS_SET_NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN_SUBTYPE:
                                           bulk penalty, no pressure DOFs */

typedef struct mife_mr_desc {
  const mife_element_desc* u; /* displacement (= geometric) basis: dim 3, dN, N, w */
  int n_pressure;             /* pressure DOFs per element: 1 (element based), the
                                 pressure basis' node count (node based), or 0
                                 (MIFE_MR_NEARLY_INCOMPRESSIBLE, which has no pressure) */
  const double* Np;           /* [n_gauss][n_pressure] pressure basis at the displacement
                                 Gauss points; NULL when n_pressure == 1 */
  int formulation;            /* MIFE_MR_DEVIATORIC, MIFE_MR_CAUCHY or
                                 MIFE_MR_NEARLY_INCOMPRESSIBLE */
} mife_mr_desc;

/* Residual, exact Jacobian and gravity vector of the Mooney-Rivlin equations,
 * accumulated (+=) into caller-zeroed global arrays.  Replaces, per element,
 * FiniteElasticity_FiniteElementResidualEvaluate (MOONEY_RIVLIN subtype
 * src/finite_elasticity_routines.F90:1669-1760 with
 * FINITE_ELASTICITY_GAUSS_STRESS_TENSOR :6985-7058, or the general branch
 * :1824-2000 with FINITE_ELASTICITY_GAUSS_CAUCHY_TENSOR :6036-6068, or the
 * NEARLY_INCOMPRESSIBLE_MOONEY_RIVLIN branch :3293-3360 with its law :5308-5320), the gravity
 * term :3370-3430, FiniteElasticity_FiniteElementJacobianEvaluate (:879-1296; here
 * the exact derivative of the residual, which iron's FD path
 * src/equations_set_routines.F90:2839-2962 approximates) and EquationsMatrices_ElementAdd.
 *   coords     [n_geom_nodes][3] reference coordinates
 *   elem_nodes [n_elem][n_nodes] geometric node of each displacement element node
 *   elem_dofs  [n_elem][3*n_nodes + n_pressure] global DOFs: x over the element
 *              nodes, then y, z, then the pressure DOFs
 *   state      [n_dofs] current dependent values (deformed coordinates, pressures)
 *   material   [n_elem][2] c1, c2; [n_elem][3] c1, c2, kappa (bulk penalty) for
 *              MIFE_MR_NEARLY_INCOMPRESSIBLE
 *   body_force NULL or host [3] = density * gravity * load fraction
 *   pos        [n_elem][ne][ne] CSR positions (ne = 3*n_nodes + n_pressure), needed
 *              with jac_vals
 *   residual, jac_vals, rhs: outputs or NULL; rhs gets the gravity vector. */
int mife_mr_assemble(const mife_mr_desc* desc, int n_elem, const double* coords,
                     const int32_t* elem_nodes, const int32_t* elem_dofs,
                     const double* state, const double* material,
                     const double* body_force, const int32_t* pos,
                     const int32_t* order, const int32_t* colour_ptr_host,
                     int n_colours, double* residual, double* jac_vals, double* rhs,
                     void* stream);

/* Staged scatter of the Mooney-Rivlin element pass.  One launch over all
 * elements writes each element's entries into its slots of a stage buffer, and
 * a fixed-order reduction adds them into the global arrays: for every CSR entry
 * k the slots e*ne*ne + a*ne + b that add into it, jac_list[jac_ptr[k] ..
 * jac_ptr[k+1]) in increasing order (element order), and for every DOF the slots
 * e*ne + a, dof_list[dof_ptr[d] .. dof_ptr[d+1]).  No reference counterpart
 * (iron adds element matrices in element order, EquationsMatrices_ElementAdd). */
typedef struct mife_mr_plan {
  int64_t nnz;              /* CSR entries (jac_ptr has nnz + 1 entries) */
  int n_dofs;               /* DOFs (dof_ptr has n_dofs + 1 entries) */
  int pad;
  const int32_t* jac_ptr;
  const int32_t* jac_list;  /* [n_elem * ne * ne] */
  const int32_t* dof_ptr;
  const int32_t* dof_list;  /* [n_elem * ne] */
} mife_mr_plan;

/* Workspace bytes of mife_mr_plan_build. */
size_t mife_mr_plan_work_bytes(int n_elem, int ne, int64_t nnz, int n_dofs);

/* Builds a plan from the scatter map pos [n_elem][ne][ne] (CSR positions) and
 * elem_dofs [n_elem][ne]: counts, an exclusive scan, an atomic fill and a sort
 * of each segment.  Synchronises the stream; fails on a position or DOF out of
 * range. */
int mife_mr_plan_build(int n_elem, int ne, const int32_t* pos, int64_t nnz, const int32_t* elem_dofs,
                       int n_dofs, int32_t* jac_ptr, int32_t* jac_list, int32_t* dof_ptr,
                       int32_t* dof_list, void* work, size_t work_bytes, void* stream);

/* Bytes of the stage buffer of mife_mr_assemble_staged: n_elem (ne^2 + 2 ne)
 * doubles, ne = 3 n_nodes + n_pressure.  Zero-fill it once before first use
 * (the pressure-pressure and gravity-pressure slots are never written). */
size_t mife_mr_stage_bytes(const mife_mr_desc* desc, int n_elem);

/* As mife_mr_assemble, with the element pass run once over all elements into
 * `stage` and the plan's fixed-order reductions (+=) into residual, jac_vals
 * and rhs.  Replaces the same reference routines as mife_mr_assemble; the sums
 * into each entry run in element order, so results do not depend on the launch
 * schedule. */
int mife_mr_assemble_staged(const mife_mr_desc* desc, int n_elem, const double* coords,
                            const int32_t* elem_nodes, const int32_t* elem_dofs,
                            const double* state, const double* material,
                            const double* body_force, const mife_mr_plan* plan, double* stage,
                            double* residual, double* jac_vals, double* rhs, void* stream);

/* --------------------------------------------------- sparse operations */

/* y = A x (prec selects FP64 or FP32 values and vectors).  For SELL, x and y
 * are in the slice numbering. */
int mife_spmv(const mife_matrix* A, const void* vals, const void* x, void* y, int prec,
              void* stream);

/* dst[k] = src[src_pos[k]] (0 where src_pos[k] < 0), src and dst each FP64 or
 * FP32 (src_prec, dst_prec).
 * The per-solve copy of K_ff into the solver matrix,
 * SOLVER_MATRIX_EQUATIONS_MATRIX_ADD (src/solver_matrices_routines.F90:806-1023). */
int mife_gather_values(int64_t nnz, const int32_t* src_pos, const void* src, int src_prec,
                       void* dst, int dst_prec, void* stream);

/* The same copy into both precisions at once, for the mixed-precision solve:
 * dst64[k] = src[src_pos[k]] and dst32[k] = (float)dst64[k] (0 where src_pos[k] < 0).
 * The FP32 inner iterations then run on the FP64 matrix rounded to FP32, read
 * from the FP64 values in the same pass (SOLVER_MATRIX_EQUATIONS_MATRIX_ADD,
 * src/solver_matrices_routines.F90:806-1023, once per precision). */
int mife_gather_values2(int64_t nnz, const int32_t* src_pos, const double* src, double* dst64,
                        float* dst32, void* stream);

/* Solver RHS over the free rows (SOLVER_MATRICES_STATIC_ASSEMBLE RHS part,
 * src/solver_routines.F90:14640-14850):
 *   b[f] = src[i] + rhs[i] - lift_i,  i = free_idx[f],
 *   lift_i = sum_{j fixed, |u_j|>=ZT, |K_ij|>=ZT} K_ij u_j for rows whose RHS DOF is
 *   free; rows with rhs_fixed[i] != 0 (a Neumann or fixed condition on the RHS
 *   variable) get no lifting, as in the reference (:14850).
 * src, rhs, rhs_fixed may be NULL.  ZT = 5*DBL_EPSILON (src/constants.F90:69). */
int mife_solver_rhs(int n_free, const int32_t* free_idx, const int32_t* row_ptr,
                    const int32_t* col, const double* vals, const int32_t* new_idx,
                    const double* u, const double* src, const double* rhs,
                    const uint8_t* rhs_fixed, double* b, void* stream);

/* u[free_idx[f]] = x[f] (SOLVER_VARIABLES_FIELD_UPDATE, src/solver_routines.F90:22814-23001). */
int mife_field_update(int n_free, const int32_t* free_idx, const double* x,
                      double* u, void* stream);

/* ------------------------------------- solver mapping with DOF constraints */

/* A constrained DOF c is tied to one free DOF m, u_c = a_c u_m
 * (BoundaryConditions_DofConstraintSet, src/boundary_condition_routines.F90:3204-3313;
 * couplings of BoundaryConditions_DofConstraintsCreateFinish, :3320-3472).  The
 * solver row / column of m stands for every equations row / column coupled to it,
 * each scaled by its coupling coefficient (SOLVER_MAPPING_CALCULATE,
 * src/solver_mapping_routines.F90:445-520, :1445-1511), so the solver matrix is
 * P^T K P:
 *   owner[g] [n]  solver DOF of equations DOF g, -1 when Dirichlet-fixed (device)
 *   coef[g]  [n]  coupling coefficient of g (1 for an unconstrained free DOF)
 * Work size for mife_couple_pattern; the same work buffer must be passed to both
 * passes. */
size_t mife_couple_pattern_work_bytes(int n, int64_t nnz, int n_s);

/* Solver pattern of P^T K P (SOLVER_MATRIX_STRUCTURE_CALCULATE,
 * src/solver_matrices_routines.F90:1612-2431): rows and columns are solver DOFs in
 * ascending order, columns sorted and unique.  An entry's value is the sum of a run
 * of equations entries (contrib_*), in ascending equations-CSR position.
 * Pass 1 (s_col == NULL): s_row_ptr[n_s+1], grp_ptr[n_s+1], grp_rows[n - n_fixed]
 * (equations rows of each solver row, ascending) and the host reads *nnz_s_out,
 * *n_contrib_out.  Pass 2: s_col[nnz_s], contrib_ptr[nnz_s+1],
 * contrib_pos[n_contrib] (equations-CSR positions), contrib_row[n_contrib]
 * (their rows).  Rejects owner values outside [-1, n_s) and solver DOFs that no
 * equations row maps to. */
int mife_couple_pattern(int n, int64_t nnz, const int32_t* row_ptr, const int32_t* col,
                        const int32_t* owner, int n_s, int32_t* grp_ptr, int32_t* grp_rows,
                        int32_t* s_row_ptr, int32_t* s_col, int32_t* contrib_ptr,
                        int32_t* contrib_pos, int32_t* contrib_row, int64_t* nnz_s_out,
                        int64_t* n_contrib_out, void* work, void* stream);

/* Solver-matrix values dst[e] = sum_run (src[p] a_row(p)) a_col(p), accumulated in
 * FP64 (SOLVER_MATRIX_EQUATIONS_MATRIX_ADD with coupling coefficients,
 * src/solver_matrices_routines.F90:935-960).  src_prec / dst_prec: MIFE_F64 or
 * MIFE_F32; col is the equations-CSR column array. */
int mife_couple_values(int64_t nnz_s, const int32_t* contrib_ptr, const int32_t* contrib_pos,
                       const int32_t* contrib_row, const int32_t* col, const double* coef,
                       const void* src, int src_prec, void* dst, int dst_prec, void* stream);

/* Solver RHS with coupled rows (SOLVER_MATRICES_STATIC_ASSEMBLE,
 * src/solver_routines.F90:14640-14850, row coupling coefficients :14660-14680):
 *   b[s] = sum_{r in group s} a_r (src[r] + rhs[r] - lift_r),
 * lift_r as in mife_solver_rhs over the Dirichlet-fixed columns (owner < 0). */
int mife_solver_rhs_coupled(int n_s, const int32_t* grp_ptr, const int32_t* grp_rows,
                            const double* coef, const int32_t* row_ptr, const int32_t* col,
                            const double* vals, const int32_t* owner, const double* u,
                            const double* src, const double* rhs, const uint8_t* rhs_fixed,
                            double* b, void* stream);

/* u[g] = a_g x[owner[g]] for every DOF with owner[g] >= 0; fixed DOFs keep their
 * values (SOLVER_VARIABLES_FIELD_UPDATE with coupling coefficients,
 * src/solver_routines.F90:22814-23001, :22877). */
int mife_field_update_coupled(int n, const int32_t* owner, const double* coef,
                              const double* x, double* u, void* stream);

/* Reactions / flux on the RHS variable, EQUATIONS_SET_BACKSUBSTITUTE CSR branch
 * (src/equations_set_routines.F90:1784-1830).  iron_compat=1 reproduces the
 * reference's u[k - rowstart] read (:1806); 0 computes the correct K u.
 * rhs_fixed (uint8 [n] or NULL) marks RHS DOFs with a fixed condition.
 * R[i] = (free ? sum : 0) - (src ? src[i] : 0). */
int mife_backsubstitute(int n, const int32_t* row_ptr, const int32_t* col,
                        const double* vals, const double* u, const double* src,
                        const uint8_t* rhs_fixed, int iron_compat, double* R,
                        void* stream);

/* ----------------------------------------------------- Krylov solvers */

/* Workspace bytes for mife_cg / mife_refine / mife_minres, and for mife_gmres
 * with a restart length up to `restart`, on n rows.  Computed from the same
 * layout the solvers take their buffers from; every solver is also given the
 * size of its workspace (work_bytes) and returns MIFE_ERR_ARG, before any
 * launch, when its layout does not fit. */
size_t mife_solver_work_bytes(int n, int restart);

/* Jacobi-preconditioned CG in the precision of `prec` (x, b and vals alike),
 * stopped when ||b - A x||_2 <= rtol ||b||_2 (recurrence residual, confirmed by
 * an explicit residual at exit) or after maxit iterations.  Replaces KSPCG+PCJACOBI
 * (SOLVER_LINEAR_ITERATIVE_CREATE_FINISH / _SOLVE, src/solver_routines.F90:10921-11157,
 * :11775-11984).  x holds the initial guess on entry.  iters and relres are
 * host outputs (the call synchronises the stream).  Returns
 * MIFE_ERR_NOT_CONVERGED when maxit is reached. */
int mife_cg(const mife_matrix* A, const void* vals, const void* b, void* x, int prec,
            double rtol, int maxit, void* work, size_t work_bytes, int* iters, double* relres,
            void* stream);

/* mife_cg with no allocation and no workspace argument: the solve's buffers
 * (two solver states of 524,544 bytes and seven vectors of n values, each
 * 256-byte aligned) live in a 64 MiB array that the library holds on each GPU it
 * runs on until the process ends, so at most 1,179,616 rows in FP64 or
 * 2,359,232 in FP32; larger systems return MIFE_ERR_ARG before any device work.
 * The call makes no hipMalloc / hipFree or other allocating call.  All calls on one
 * GPU share that array and it keeps the last call's contents (every buffer is
 * written before it is read): calls that overlap on one GPU, from several host
 * threads or streams, race and must be serialised by the caller.  Results are those
 * of mife_cg.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid hipMalloc). */
int mife_cg_noalloc(const mife_matrix* A, const void* vals, const void* b, void* x, int prec,
                    double rtol, int maxit, int* iters, double* relres, void* stream);

/* Mixed-precision solve: Jacobi CG on the FP32 operator vals32 whose residual is
 * replaced by the FP64 residual r = b - A64 x each time the FP32 recurrence
 * residual has fallen by the factor inner_rtol since the last replacement
 * ("reliable updates"); the FP32 correction is folded into the FP64 solution x
 * at each replacement, and the search direction is kept, so FP64 accuracy is
 * won back without restarting the Krylov space.  Stops when
 * ||b - A64 x|| <= rtol ||b||.  iters = FP32 iterations, outer = FP64
 * replacements.  No reference counterpart: iron solves in FP64 only
 * (Petsc_KSPSolve, src/solver_routines.F90:11869). */
int mife_refine(const mife_matrix* A, const double* vals64, const float* vals32,
                const double* b, double* x, double rtol, double inner_rtol, int maxit,
                void* work, size_t work_bytes, int* iters, int* outer, double* relres,
                void* stream);

/* Right-preconditioned restarted GMRES(restart) in FP64 (KSPGMRES default
 * restart 30, src/solver_routines.F90:11320-11369), restart <= 256 (workspace
 * from mife_solver_work_bytes(n, restart)).  The
 * preconditioner is diagonal: Jacobi 1/a_ii, except rows where pc_inv[i] != 0
 * (device, caller's row order, may be NULL) which use pc_inv[i] instead, e.g. a
 * scaled inverse pressure mass for the zero-diagonal rows of a mixed
 * displacement/pressure Jacobian (PCJACOBI would use 1 there). */
int mife_gmres(const mife_matrix* A, const double* vals, const double* b, double* x,
               const double* pc_inv, double rtol, int restart, int maxit, void* work,
               size_t work_bytes, int* iters, double* relres, void* stream);

/* Preconditioned MINRES (Paige-Saunders) for symmetric, possibly indefinite A,
 * such as the Newton Jacobian of the Mooney-Rivlin deviatoric formulation.
 * Vectors and scalars are FP64; vals are FP64 or FP32 (prec), FP32 values being
 * read into FP64 arithmetic, which halves the value stream of every product.
 * SPD diagonal preconditioner |1/a_ii|, or pc_inv[i] where nonzero (as for
 * mife_gmres, absolute value taken).  Stops when the preconditioned residual
 * norm has fallen by rtol or after maxit iterations; relres returns the true
 * ||b - A x|| / ||b|| of the operator given, at exit.  No reference counterpart
 * (iron's Newton solves go through KSP, src/solver_routines.F90:18674-18763).
 * Workspace: mife_solver_work_bytes(n, 1). */
int mife_minres(const mife_matrix* A, const void* vals, int prec, const double* b, double* x,
                const double* pc_inv, double rtol, int maxit, void* work, size_t work_bytes,
                int* iters, double* relres, void* stream);

/* -------------------------------------------- grouped solves (serving) */

/* One independent system of a group, stored in shared pools (mife_pools) with
 * local 0-based columns.  CSR: n+1 row pointers at rp_off, columns and values at
 * nz_off.  SELL-64 (as mife_matrix): n_slices+1 slice pointers at rp_off,
 * values at nz_off, columns at nz_off of the int32 pool or, with cb_off >= 0,
 * 16-bit offsets at nz_off of the 16-bit pool plus n_slices base columns at
 * cb_off; the system's vectors are then in slice numbering.  Vectors live at
 * row_off of the vector pools.  slot (0 <= slot < max_segments) names the
 * segment's persistent CG state.  prec: MIFE_F64 (FP64 CG) or MIFE_F32 (FP32 CG
 * with FP64 reliable updates).  The library sees the pools as bare pointers: the
 * caller keeps every segment's ranges (rows, pointers, entries, base columns)
 * inside its allocations. */
typedef struct mife_segment {
  int32_t n;
  int32_t maxit;
  int32_t slot;
  int32_t format;   /* MIFE_FMT_CSR, MIFE_FMT_SELL or MIFE_FMT_SELL3 (16-bit block columns at col16 + nz_off) */
  int64_t row_off;
  int64_t rp_off;   /* CSR: row_ptr[n+1]; SELL: slice_ptr[n_slices+1] (entry offsets, multiples of 64) */
  int64_t nz_off;
  double rtol;
  int32_t n_slices; /* SELL: ceil(n / 64) */
  int32_t prec;     /* MIFE_F64 or MIFE_F32 */
  int64_t cb_off;   /* SELL: offset of the per-slice base columns (16-bit columns), or -1 (int32 columns);
                       SELL3: required */
} mife_segment;

/* Shared storage of the grouped segments: row / slice pointers, int32 columns,
 * 16-bit slice-relative columns with their per-slice base columns, FP64 values
 * and their FP32 copy at the same offsets (read by MIFE_F32 segments only). */
typedef struct mife_pools {
  const int32_t* rp;
  const int32_t* col;
  const uint16_t* col16;
  const int32_t* cbase;
  const double* val;
  const float* val32;
} mife_pools;

/* Grouped Jacobi CG: every launch serves all segments of a table; each segment
 * has its own scalars, its own fixed-order reductions and its own stopping
 * test, so a system's iterates do not depend on which other systems share the
 * launches.  No reference counterpart (iron solves one system per KSPSolve,
 * src/solver_routines.F90:11869); it is the batching layer of the serving path.
 * Per segment precision:
 *   MIFE_F64: FP64 CG, done when the recurrence residual ||r|| <= rtol ||b||.
 *   MIFE_F32: FP32 CG on the FP32 values with FP64 reliable updates (the
 *     algorithm of mife_refine): every check_every iterations (and after the
 *     last of a call) a segment whose FP32 r.r has fallen by inner_rtol^2 since
 *     its last replacement (or below a quarter of its target, or at maxit) folds
 *     x += x32 and replaces r by the FP64 residual b - A64 x; it is done when
 *     that true residual meets rtol (or at maxit).
 *   mife_grouped_state_bytes / _work_bytes: device buffers for the per-slot
 *     state and for one table (max_blocks from the sum of mife_grouped_blocks(n)).
 *   mife_grouped_setup: builds the table of n_seg segments (host staging buffer
 *     of _work_bytes), synchronises the stream.
 *   mife_grouped_cg_init: Jacobi diagonal, r = b - A x (FP64), z, p = q = 0 (and
 *     x32 = 0) and the state of the table's segments.
 *   mife_grouped_cg_iterate: `iterations` iterations of every unconverged
 *     segment; launches only (no allocation, memset or synchronisation), so the
 *     call can be captured in a HIP graph and replayed.
 *   mife_grouped_cg_residual: true ||b - A x|| of the table's segments.
 *   mife_grouped_cg_status: host copy of iterations, done flags and relative
 *     residuals of all max_segments slots (synchronises); the relative residual
 *     is the last true one where one was computed, else the recurrence's.
 * Vector pools: b, x [pool_rows] (FP64; x holds the solution), vec5
 * [5][pool_rows] (FP64 D^-1, r, z, p, q of MIFE_F64 segments) and vec6
 * [6][pool_rows] (FP32 D^-1, r, z, p, q, x32 of MIFE_F32 segments). */
size_t mife_grouped_state_bytes(int max_segments);
size_t mife_grouped_work_bytes(int max_segments, int max_blocks);
int mife_grouped_blocks(int n_rows);
int mife_grouped_setup(const mife_segment* segs_host, int n_seg, int max_segments, int max_blocks,
                       int* n_blocks_out, void* table_host, void* work, void* stream);
int mife_grouped_cg_init(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                         const double* b, const double* x, double* vec5, float* vec6, int64_t pool_rows,
                         void* work, void* state, void* stream);
int mife_grouped_cg_iterate(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                            const double* b, double* x, double* vec5, float* vec6, int64_t pool_rows,
                            int iterations, int check_every, double inner_rtol, void* work, void* state,
                            void* stream);
int mife_grouped_cg_residual(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                             const double* b, const double* x, void* work, void* state, void* stream);
int mife_grouped_cg_status(int max_segments, const void* state, int* iters, int* done, double* relres,
                           void* stream);

/* ------------------------------------------------------------- profiling */

/* Kernel timing with HIP events recorded on the launching stream around each
 * hot kernel (CG SpMV / update, assembly colour passes, FP64 residual).  Off by
 * default; no reference counterpart (replaces the TAUPROF / timing-output
 * instrumentation, src/equations_set_routines.F90:1239-1353).  mife_profile_read
 * synchronises the recorded events and returns per-kernel launch counts and
 * total milliseconds; names are written as n_max slots of 48 chars. */
int mife_profile_enable(int on);
int mife_profile_reset(void);
int mife_profile_read(int n_max, char* names, int* counts, double* total_ms, int* n_out);

/* ---------------------------------------------------------- HIP graphs */

/* CG solves (mife_cg, mife_refine) on systems of at most 131,072 rows replay
 * their iterations from a cached HIP graph of 8 iterations (captured on first
 * use for each set of kernel arguments) when profiling is off; on by default.
 * The kernels and their order are those of the host launches, so results are
 * bit for bit the same.  mife_cg_noalloc never takes this path (capture and
 * instantiation may allocate).  No reference counterpart (PETSc's KSPSolve
 * launches nothing). */
int mife_graphs_enable(int on);
/* Number of instantiated CG chunk graphs held by the library (at most 16). */
int mife_graphs_cached(void);

/* Classical Gram-Schmidt refinement of mife_gmres, process-wide: 0 never (one
 * pass per iteration), 1 always (a second pass with the correction).  Mirrors
 * PETSc's KSPGMRESSetCGSRefinementType (KSP_GMRES_CGS_REFINE_NEVER /
 * _ALWAYS), which iron leaves at PETSc's default (src/solver_routines.F90
 * sets only the restart, via Petsc_KSPGMRESSetRestart, src/cmiss_petsc.F90:2659). */
int mife_gmres_cgs_refinement(int mode);

#ifdef __cplusplus
}
#endif

#endif /* MIFE_H */
