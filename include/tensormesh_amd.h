/*
 * tensormesh_amd.h -- C-ABI of libtmhip.so, the MI355X (gfx950) kernels of the
 * tensorized element-assembly path.
 *
 * The reference (kilickursat/TensorMesh) is pure Python/PyTorch and has no
 * FFI; its "operator API" is the Python classes ElementAssembler,
 * NodeAssembler, ReduceProjector, Condenser and SparseMatrix.  Each entry
 * point below replaces one step of that path and cites the reference lines
 * it stands for.  The Python package tensormesh_amd binds these with ctypes
 * (tensormesh_amd/_native.py).
 *
 * Conventions
 *  - Device pointers unless a parameter says "host".  All memory is owned by
 *    the caller; no entry point allocates, frees or synchronises, except the
 *    host-side pattern builders, which run on the calling thread.
 *  - Every device entry point is stream-ordered on the hipStream_t it gets.
 *  - Return 0 on success, a negative tm_status otherwise; tm_last_error()
 *    gives a thread-local message.  No C++ exception crosses the ABI.
 *  - Indices are int32; entry points reject sizes that do not fit.
 *  - dtype TM_F64: points, coefficients and values are double.  TM_F32: they
 *    are float; the P1 scalar kernels evaluate element geometry in float and
 *    redo slivers in double (TM_FLAG_F32_NO_REDO), the other kernels evaluate
 *    geometry in double; contraction and accumulation are in float.
 */
#ifndef TENSORMESH_AMD_H
#define TENSORMESH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* tm_stream_t; /* == hipStream_t */

#define TM_ABI_VERSION 3  /* 2: tm_pattern_view gained entry / entry_t; 3: tm_rows_view gained dest / ent_start */

enum tm_status {
  TM_OK = 0,
  TM_ERR_ARG = -1,         /* bad pointer / shape / enum value */
  TM_ERR_SIZE = -2,        /* size beyond int32 indexing or a kernel limit */
  TM_ERR_HIP = -3,         /* HIP runtime error (launch / memcpy) */
  TM_ERR_UNSUPPORTED = -4, /* element / form / dtype combination not built */
  TM_ERR_TOPOLOGY = -5     /* repeated node inside one element */
};

enum tm_dtype { TM_F32 = 0, TM_F64 = 1 };

enum tm_cell {
  TM_TRI3 = 0,  /* "triangle"  B=3  D=2 */
  TM_TET4 = 1,  /* "tetra"     B=4  D=3 */
  TM_TRI6 = 2,  /* "triangle6" B=6  D=2 */
  TM_TET10 = 3, /* "tetra10"   B=10 D=3 */
  /* tensor-product cells on [0,1]^D, lexicographic vertices, then edge,
   * face and cell nodes (element.py:1801-1935, 2090-2222) */
  TM_QUAD4 = 4, /* "quad"         B=4  D=2 */
  TM_QUAD9 = 5, /* "quad9"        B=9  D=2 */
  TM_HEX8 = 6,  /* "hexahedron"   B=8  D=3 */
  TM_HEX27 = 7, /* "hexahedron27" B=27 D=3 */
  /* cubic simplices: vertices, two nodes per edge (at 1/3 and 2/3 from the
   * edge's first vertex), one centroid node per face (basis.py:639-859) */
  TM_TRI10 = 8, /* "triangle10"   B=10 D=2 */
  TM_TET20 = 9  /* "tetra20"      B=20 D=3 */
};

enum tm_form {
  TM_FORM_LAPLACE = 0,        /* gradu @ gradv                 builtin.py:61-62 */
  TM_FORM_DIFFUSION_ELEM = 1, /* kappa[e] * (gradu @ gradv)    thermal_topology.py:48-52 */
  TM_FORM_DIFFUSION_NODE = 2, /* kappa_h(x) * (gradu @ gradv)  forms.rst:172-184 */
  TM_FORM_MASS = 3,           /* u * v                         builtin.py:99-100 */
  TM_FORM_LINEAR_ELASTICITY = 4 /* Ba^T C Bb (3x3 blocks)      builtin.py:172-178 */
};

/* Element block on the device. */
typedef struct tm_mesh_view {
  int32_t cell;          /* enum tm_cell */
  int32_t dtype;         /* enum tm_dtype of points / coefficients / values */
  int64_t n_elem;
  int64_t n_points;
  const int32_t* conn;   /* [n_elem, B] row-major */
  const void* points;    /* [n_points, D] row-major */
} tm_mesh_view;

/*
 * Reference-element tables, float64, for Q quadrature points, B basis
 * functions and D dimensions:
 *   [0, Q)                          w[q]
 *   [Q, Q + Q*B)                    phi[q][b]
 *   [Q + Q*B, Q + Q*B + Q*D*B)      dphi/dxi[q][d][b]
 */
typedef struct tm_form_desc {
  int32_t form;              /* enum tm_form */
  int32_t n_quad;            /* Q */
  const double* tab_host;    /* host copy of the tables (read during the call) */
  const double* tab_dev;     /* device copy (read by P2 kernels) */
  const void* coeff;         /* DIFFUSION_ELEM: [n_elem]; DIFFUSION_NODE: [n_points]; else NULL */
  double young;              /* LINEAR_ELASTICITY: E  */
  double poisson;            /* LINEAR_ELASTICITY: nu */
  int32_t flags;             /* TM_FLAG_* */
  int32_t epw;               /* elements per wavefront of the wave kernels, 0 = the kernel's own choice:
                                element matrices 1 or 2, Neo-Hookean 1, 2 or 4 (VALU tangent), J2 1, 4
                                or 8 (cells of <= 4 nodes with rules of <= 4 points); a count the cell,
                                rule or LDS budget does not admit is TM_ERR_ARG.  Results do not depend
                                on it (one lane group per element, the same arithmetic). */
} tm_form_desc;

/* tm_form_desc.flags */
#define TM_FLAG_MFMA 1       /* float32 block tangents: contract on v_mfma_f32_32x32x2f32 */
#define TM_FLAG_UPPER 2      /* symmetric forms on the wave, Neo-Hookean and J2 kernels:
                                compute and store only the element blocks (i, j) with
                                i <= j, packed: block p (row-major over the pairs i <= j)
                                of element e at slot e*B(B+1)/2 + p, so slot workspaces
                                need n_elem*B(B+1)/2*hh values; the pattern's seg_slot
                                must then be the list of tm_symmetric_segments.  The
                                reference stores every block (element_assembler.py:
                                811-830); the assembled values are the same.  Rejected
                                (TM_ERR_ARG) by the P1 scalar kernels, the MFMA tangent
                                and the block-row path. */
#define TM_FLAG_F32_NO_REDO 4 /* diagnostic, P1 scalar kernels on float32 meshes: keep the
                                float32 geometry of every element.  By default an element
                                whose shape quality q = |det J| / prod_k |x_{k+1} - x_0|
                                is below 0.1 (a sliver, where float32 cancellation in
                                det J and the cofactors exceeds ~1e-6 of the element's
                                largest entry) is evaluated again in float64.  Float64
                                meshes are always evaluated in float64. */

/* Sparsity pattern plus the slot -> nnz permutation in segment form. */
typedef struct tm_pattern_view {
  int64_t nnz;               /* node-pair entries */
  const int32_t* row_ptr;    /* [n_points + 1] CSR */
  const int32_t* col_idx;    /* [nnz] */
  const int32_t* seg_ptr;    /* [nnz + 1] start of each entry's contributions */
  const int32_t* seg_slot;   /* [n_elem*B*B] slot ids s = e*B*B + i*B + j, ascending per entry */
  int64_t n_blocks;          /* reduction plan (tm_reduce_plan), or 0 for one thread per entry */
  const int32_t* blk_ptr;    /* [n_blocks + 1] first entry of each reduction block */
  /* Optional paired form (both NULL: segment k is entry k).  With them, nnz
   * counts the listed segments, segment k is written to entry[k] and,
   * transposed, to entry_t[k] (tm_segmented_reduce_pairs); outputs keep the
   * pattern's full entry count.  For symmetric forms: the upper entries of a
   * TM_FLAG_UPPER segment list fill the lower ones. */
  const int32_t* entry;      /* [nnz] */
  const int32_t* entry_t;    /* [nnz]; == entry[k] on the diagonal */
  /* Optional contiguous contributions (tm_slot_dest_build), with a
   * TM_FLAG_UPPER segment list: the element kernels of tm_assemble_bilinear,
   * tm_assemble_neohookean and tm_assemble_j2 launched with TM_FLAG_UPPER
   * write packed block p of element e at slot (slot_dest[e*P + p] &
   * INT32_MAX), P = B(B+1)/2, transposed when bit 31 is set; segment k's
   * contributions are then the consecutive slots [seg_ptr[k], seg_ptr[k+1])
   * and seg_slot is not read.  Same values and summation order. */
  const int32_t* slot_dest;  /* [n_elem*P] or NULL */
} tm_pattern_view;

/* Node -> (element, local index) incidence, ascending element id per node. */
typedef struct tm_incidence_view {
  const int32_t* inc_ptr;    /* [n_points + 1] */
  const int32_t* inc;        /* [n_elem*B] packed e*B + i */
  int64_t n_blocks;          /* reduction plan over nodes, or 0 */
  const int32_t* blk_ptr;    /* [n_blocks + 1] */
} tm_incidence_view;

/* Row-side view of a P1 pattern for the fused path: CSR rows, node
 * incidence, and per incidence the packed positions (byte j = position of
 * conn[e,j] inside row conn[e,i]). */
typedef struct tm_rows_view {
  int64_t nnz;
  const int32_t* row_ptr;    /* [n_points + 1] */
  const int32_t* inc_ptr;    /* [n_points + 1] */
  const int32_t* inc;        /* [n_elem*B] packed e*B + i */
  const uint32_t* colpos;    /* [n_elem*B] */
  const uint32_t* entry_mask; /* [nnz] bit t: the row's t-th incidence contributes (rows with <= 32 incidences), or NULL */
  int64_t n_blocks;          /* reduction plan over inc_ptr (rows) */
  const int32_t* blk_ptr;    /* [n_blocks + 1] */
  const int32_t* blk_info;   /* [n_blocks + 1][4] {blk_ptr[b], inc_ptr[blk_ptr[b]], row_ptr[blk_ptr[b]], 0},
                                or NULL (then read through blk_ptr) */
  /* Optional contribution order (tm_rows_dest_build), both or neither: each
   * workgroup lists its rows' contributions entry by entry, ascending
   * element within an entry; dest gives, per incidence, the list position of
   * its B values (natural vertex order, uint16 each, 4 per incidence) and
   * ent_start, per entry, where its contributions begin in its workgroup's
   * list.  With them an entry is a contiguous sum (same order, same values). */
  const uint16_t* dest;      /* [n_elem*B][4] */
  const uint16_t* ent_start; /* [nnz] */
} tm_rows_view;

/* Block rows of a node-pair pattern for the block-row reduction: runs of
 * whole rows (nodes) per workgroup, and for every contribution (in seg_ptr /
 * seg_slot order: ascending slot within each entry) its workgroup-local
 * (incidence, column) pair. */
typedef struct tm_block_rows_view {
  int64_t nnz;               /* block entries */
  const int32_t* row_ptr;    /* [n_points + 1] */
  const int32_t* seg_ptr;    /* [nnz + 1] contributions per entry */
  const uint16_t* pairs;     /* [n_elem*B*B] from tm_block_rows_pairs */
  const int32_t* inc_ptr;    /* [n_points + 1] node incidence lists (ascending e*B + i) */
  const int32_t* inc;        /* [n_elem*B] */
  int64_t n_blocks;
  const int32_t* blk_info;   /* [n_blocks + 1][4] {first row r, inc_ptr[r], row_ptr[r], seg_ptr[row_ptr[r]]};
                                at most 256 / D rows and 6553 incidences per block; blocks of more
                                than 1024 / D^2 entries take several passes */
} tm_block_rows_view;

/* Selected boundary facets (FacetAssembler): facet i is local facet local[i]
 * of element elem[i] (triangle: edges (1,2), (0,2), (0,1); tetrahedron: faces
 * (1,2,3), (0,2,3), (0,1,3), (0,1,2); quadrilateral: edges (0,1), (0,2), (1,3),
 * (2,3); hexahedron: faces (0,1,2,3), (0,1,4,5), (0,2,4,6), (1,3,5,7),
 * (2,3,6,7), (4,5,6,7)), in the reference's selection order. */
typedef struct tm_facet_view {
  int64_t n_sel;
  const int32_t* elem;       /* [n_sel] */
  const int32_t* local;      /* [n_sel] */
} tm_facet_view;

/* Block-diagonal stack of SPD CSR systems for the batched Jacobi-PCG.
 * Workgroups are system-aligned: workgroup w covers rows
 * [blk_row0[w], blk_row1[w]) of system blk_sys[w] (at most 256 rows);
 * sys_blk[s]..sys_blk[s+1] are system s's workgroups. */
typedef struct tm_pcg_view {
  int32_t dtype;             /* enum tm_dtype of values and vectors */
  int32_t reserved;
  int64_t n_rows;            /* total rows of all systems */
  int64_t n_sys;
  int64_t n_blocks;
  const int32_t* row_ptr;    /* [n_rows + 1] global CSR (block diagonal) */
  const int32_t* col_idx;    /* [nnz] global column ids */
  const void* values;        /* [nnz] */
  const void* dinv;          /* [n_rows] inverse diagonal (tm_jacobi) */
  void* r;                   /* [n_rows] workspaces */
  void* z;
  void* p;
  void* q;
  double* partials;          /* [n_blocks * 3] */
  double* scalars;           /* [n_sys * 8]: rz, pq, alpha, beta, |r|^2, |b|^2, active, iterations */
  const int32_t* blk_sys;    /* [n_blocks] */
  const int32_t* blk_row0;   /* [n_blocks] */
  const int32_t* blk_row1;   /* [n_blocks] */
  const int32_t* sys_blk;    /* [n_sys + 1] */
  double rtol;               /* stop system s once |r_s| <= rtol |b_s| */
  void* s;                   /* [n_rows] A p (two- and one-launch variants) */
  int32_t two_launch;        /* launches per iteration: 0 five (partials [3 n_blocks], scalars [8 n_sys]),
                                1 two (partials [5 n_blocks], scalars [16 n_sys]),
                                2 one (partials [7 n_blocks], scalars [16 n_sys], aux) */
  int32_t parity;            /* two- and one-launch: iterations run since tm_pcg_init, mod 2 */
  void* aux;                 /* one-launch: [2 n_rows], the second copies of A u and A p */
} tm_pcg_view;

int tm_abi_version(void);
const char* tm_last_error(void);

/* Workspace sizes, in bytes, for a mesh (host arithmetic, no device call).
 *   TM_WS_ELEMENT_SLOTS: slot_ws of tm_assemble_bilinear / tm_element_matrices
 *     and tslot_ws of the Neo-Hookean and J2 tangents: n_elem*B*B*hh values
 *     (hh = 1 scalar forms, D*D block forms)
 *   TM_WS_NODE_SLOTS: slot_ws of tm_assemble_linear (hh = 1) and rslot_ws of
 *     the Neo-Hookean / J2 residuals (hh = D): n_elem*B*hh values
 *   TM_WS_ENERGY: energy_ws of tm_assemble_j2: n_elem + TM_J2_ENERGY_PARTS
 *     doubles (hh ignored)
 *   TM_WS_ELEMENT_SLOTS_UPPER: as TM_WS_ELEMENT_SLOTS with TM_FLAG_UPPER:
 *     n_elem*B(B+1)/2*hh values
 * Values are of the mesh dtype.  Returns the size, or a negative tm_status.
 * The reference sizes these buffers inside torch ops (no counterpart). */
enum tm_workspace { TM_WS_ELEMENT_SLOTS = 0, TM_WS_NODE_SLOTS = 1, TM_WS_ENERGY = 2, TM_WS_ELEMENT_SLOTS_UPPER = 3 };
int64_t tm_workspace_bytes(int32_t op, const tm_mesh_view* mesh, int32_t hh);

/* ---------------------------------------------------------------------
 * Pattern (host).  Replaces the edge / projector construction of
 * ElementAssembler.from_mesh (tensormesh/assemble/element_assembler.py:804-853):
 * CSR of the sorted unique (conn[e,i], conn[e,j]) pairs -- the order scipy
 * coo->csr->coo gives -- and, per entry, the ascending list of slots
 * s = e*B*B + i*B + j that ReduceProjector.index_add_ would sum into it.
 * All pointers are HOST pointers.  tm_pattern_nnz sizes the outputs;
 * tm_pattern_build fills them (slot2nnz may be NULL).
 * ------------------------------------------------------------------- */
int tm_pattern_nnz(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                   int64_t* nnz_out);
int tm_pattern_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                     int32_t* row_ptr, int32_t* col_idx, int32_t* seg_ptr, int32_t* seg_slot,
                     int32_t* slot2nnz);

/* Reduction plan (host): split entries [0, n_entries) into blocks of whole
 * entries holding at most `cap` contributions each (an entry with more gets a
 * block of its own).  blk_ptr_out needs n_entries + 1 slots; *n_blocks gets
 * the count.  Used by the segmented reduction to stage each block's
 * contributions in LDS. */
int tm_reduce_plan(const int32_t* seg_ptr, int64_t n_entries, int32_t cap, int32_t* blk_ptr_out,
                   int64_t* n_blocks);

/* Packed column positions of every incidence (host), for tm_rows_view, and
 * optionally (entry_mask != NULL) the per-entry incidence masks.
 * B <= 4 and rows of at most 255 entries. */
int tm_colpos_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                    const int32_t* row_ptr, const int32_t* col_idx, const int32_t* inc_ptr, const int32_t* inc,
                    uint32_t* colpos, uint32_t* entry_mask);

/* Workgroup-local (incidence, column) pair of every contribution (device,
 * stream-ordered): for contribution c of entry k of row r, with slot
 * s = seg_slot[c] = e*B*B + a*B + b, pairs[c] = (inc_inv[e*B + a] - row_t0[r])*B + b,
 * where row_t0[r] is the first incidence of r's block.  *overflow is set to 1
 * (device int32, caller-zeroed) if a pair does not fit 16 bits. */
int tm_block_rows_pairs(int32_t n_basis, int64_t n_points, const int32_t* row_ptr, const int32_t* seg_ptr,
                        const int32_t* seg_slot, const int32_t* inc_inv, const int32_t* row_t0, uint16_t* pairs,
                        int32_t* overflow, tm_stream_t stream);

/* Device pattern builder (stream-ordered; all pointers device).  Same arrays as
 * tm_incidence_build + tm_pattern_build + tm_colpos_build.
 *   tm_pattern_device_count: inc_ptr [N+1], inc [E*B], row_ptr [N+1]; work is
 *     int32 scratch of N+1 entries; afterwards nnz = row_ptr[N] and work[N] != 0
 *     flags a row with more than 2048 neighbour keys (then use the host builder).
 *   tm_pattern_device_fill: col_idx [nnz], seg_ptr [nnz+1], seg_slot and
 *     slot2nnz [E*B*B], colpos [E*B] and entry_mask [nnz] (B <= 4, may be NULL). */
int tm_pattern_device_count(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                            int32_t* inc_ptr, int32_t* inc, int32_t* row_ptr, int32_t* work, tm_stream_t stream);
int tm_pattern_device_fill(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                           const int32_t* inc_ptr, const int32_t* inc, const int32_t* row_ptr, int32_t* col_idx,
                           int32_t* seg_ptr, int32_t* seg_slot, int32_t* slot2nnz, uint32_t* colpos,
                           uint32_t* entry_mask, tm_stream_t stream);

/* Node incidence (host), the scatter map of NodeAssembler's ReduceProjector
 * (tensormesh/assemble/node_assembler.py:665-670) in gather form. */
int tm_incidence_build(const int32_t* conn, int64_t n_elem, int32_t n_basis, int64_t n_points,
                       int32_t* inc_ptr, int32_t* inc);

/* ---------------------------------------------------------------------
 * Bilinear assembly, two fixed stages on one stream.
 * Replaces ElementAssembler.__call__ for the built-in forms
 * (element_assembler.py:231-445): geometry Transformation.batch_shape_grad_jxw
 * (transformation.py:1219-1276, element.py:874-968), form + _integrate
 * (element_assembler.py:190-209, 416-438), ReduceProjector.__call__
 * (projector.py:95-123).
 *   stage G: per element -> slot_ws[n_elem*B*B*H*H] (H = 3 for elasticity, else 1)
 *   stage S: segmented reduction slot_ws -> values[nnz*H*H] in entry order,
 *            each entry summed in ascending slot order (no atomics).
 * values is the edata of SparseMatrix (scalar) or the block data handed to
 * SparseMatrix.from_block_coo (matrix.py:331-381) for H > 1.
 * ------------------------------------------------------------------- */
int tm_assemble_bilinear(const tm_mesh_view* mesh, const tm_form_desc* form,
                         const tm_pattern_view* pat, void* slot_ws, void* values,
                         tm_stream_t stream);

/* Fused row-centric P1 assembly (scalar forms on TRI3 / TET4): same values,
 * same entry order and the same per-entry summation order as
 * tm_assemble_bilinear, without the slot buffer -- every CSR row walks its
 * incident elements in ascending order and recomputes their local rows.
 * Replaces the same reference chain as tm_assemble_bilinear. */
int tm_assemble_p1_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_rows_view* rows,
                        void* values, tm_stream_t stream);

/* Contribution order of the fused P1 path (dest / ent_start of tm_rows_view)
 * from a rows view with colpos and blk_info set: one workgroup per plan
 * block; blocks over the kernel's staging caps (512 incidences, 64 rows,
 * 1024 entries) are skipped (the fused kernel walks colpos for them).
 * Device pointers; dest [n_elem*B][4] uint16 (unused lanes of triangles are
 * 0), ent_start [nnz] uint16.  No reference counterpart: it is the
 * ReduceProjector permutation (projector.py:95-123) arranged per workgroup. */
int tm_rows_dest_build(int32_t n_basis, const tm_rows_view* rows, uint16_t* dest, uint16_t* ent_start,
                       tm_stream_t stream);

/* Adjoint of the P1 scalar assembly (TRI3 / TET4; laplace, diffusion, mass).
 * Given grad_values = dL/dvalues [nnz] and slot2nnz [n_elem*B*B] (the slot ->
 * entry map of tm_pattern_build), writes
 *   grad_coeff  [n_elem] (DIFFUSION_ELEM) or [n_points] (DIFFUSION_NODE), may be NULL
 *   grad_points [n_points*D], may be NULL
 * Workspaces: ws_points [n_elem*B*D] when grad_points is wanted, ws_coeff
 * [n_elem*B] for a nodal coefficient; inc (node incidence) for both node
 * reductions.  Replaces autograd through the reference's assembly
 * (element_assembler.py:416-441, element.py:956-968, transformation.py:1265-1272)
 * driven by the sparse solve's dL/dA_ij = -lambda_i u_j (torch_solve.py:262-272). */
int tm_assemble_p1_adjoint(const tm_mesh_view* mesh, const tm_form_desc* form, const int32_t* slot2nnz,
                           const void* grad_values, const tm_incidence_view* inc, void* ws_points,
                           void* ws_coeff, void* grad_points, void* grad_coeff, tm_stream_t stream);

/* Neo-Hookean residual and consistent tangent at displacement disp [n_points, D].
 * psi = mu/2 (I1 - d) - mu ln J + lam/2 (ln J)^2, mu and lam from form->young /
 * form->poisson (builtin.py:280-314); the reference obtains R and K by autograd of
 * NeoHookeanModel.energy (element_assembler.py:447-659).
 *   tangent  [nnz*D*D]     block data in node-pair CSR order (as LINEAR_ELASTICITY),
 *                          needs pat and tslot_ws [n_elem*B*B*D*D]; may be NULL
 *   residual [n_points*D]  dE/du, needs inc and rslot_ws [n_elem*B*D]; may be NULL */
int tm_assemble_neohookean(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp,
                           const tm_pattern_view* pat, const tm_incidence_view* inc, void* tslot_ws,
                           void* tangent, void* rslot_ws, void* residual, tm_stream_t stream);

/* Neo-Hookean energy at displacement disp [n_points, D]: replaces
 * NeoHookeanModel.energy (builtin.py:316-333) through ElementAssembler.energy
 * (element_assembler.py:447-659), energy = sum over elements and quadrature
 * points of psi(F) * jxw with psi as for tm_assemble_neohookean.
 *   energy [1]            mesh dtype; summed in a fixed order (each element's
 *                         points in order, then TM_J2_ENERGY_PARTS chunk
 *                         partials, then one workgroup)
 *   energy_ws             double [n_elem + TM_J2_ENERGY_PARTS]; receives the
 *                         per-element energies [n_elem]
 *   residual [n_points*D] dE/du from the same element pass when non-NULL
 *                         (the reference's energy().backward()); needs inc
 *                         and rslot_ws [n_elem*B*D]
 * form->flags is ignored (no tangent is formed). */
int tm_neohookean_energy(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp, void* energy_ws,
                         void* energy, const tm_incidence_view* inc, void* rslot_ws, void* residual,
                         tm_stream_t stream);

/* Neo-Hookean residual and tangent through incidence-ordered slot rows: the
 * element kernel writes the block row (e, a) (B blocks of D x D values, and D
 * residual values) at position inc_inv[e*B + a] of the node-sorted incidence
 * list, then a block-row reduction sums every entry's contributions in
 * ascending element order (as tm_assemble_neohookean, whose segment
 * reduction gathers the same values slot by slot).  Outputs as
 * tm_assemble_neohookean: tangent [nnz*D*D] in BSR(D x D) node-pair CSR
 * order, residual [n_points*D]; either may be NULL.  tslot_ws: [n_elem*B*B*D*D],
 * rslot_ws: [n_elem*B*D], in the mesh dtype.  TM_FLAG_MFMA selects the
 * matrix-core tangent (float32); TM_FLAG_UPPER is rejected (the block rows
 * are read whole).  Replaces the reference's autograd residual /
 * Hessian of NeoHookeanModel.energy (builtin.py:219-333,
 * element_assembler.py:447-659) with the ReduceProjector sum
 * (projector.py:95-123). */
int tm_assemble_neohookean_rows(const tm_mesh_view* mesh, const tm_form_desc* form, const void* disp,
                                const tm_block_rows_view* rows, const int32_t* inc_inv, void* tslot_ws,
                                void* tangent, void* rslot_ws, void* residual, tm_stream_t stream);

/* ---------------------------------------------------------------------
 * J2 (von Mises) plasticity with linear isotropic hardening.
 * Replaces J2Plasticity (tensormesh/assemble/builtin.py:335-561) through
 * ElementAssembler.energy (element_assembler.py:447-659): the algorithmic
 * incremental potential
 *   psi = K/2 tr(eps_tr)^2 + mu |dev eps_tr|^2 - (2 mu + 2H/3) dg^2 / 2,
 *   eps_tr = sym(grad u) (3 x 3, zero-padded in 2D) - eps_p_n,
 *   dg = max(2 mu |dev eps_tr| - sqrt(2/3)(sig0 + H alpha_n), 0) / (2 mu + 2H/3),
 * its gradient (the reference's autograd residual), its consistent Hessian
 * and the radial-return update of update_state (builtin.py:513-561).
 * ------------------------------------------------------------------- */
typedef struct tm_j2_material {
  double young;              /* E */
  double poisson;            /* nu */
  double sig0;               /* initial yield stress */
  double hardening;          /* H, linear isotropic hardening modulus */
} tm_j2_material;

/* Per-quadrature-point history, in the mesh dtype (J2Plasticity.history). */
typedef struct tm_j2_state {
  void* eps_p;               /* [n_elem][n_quad][3][3] plastic strain eps_p_n */
  void* alpha;               /* [n_elem][n_quad] equivalent plastic strain alpha_n */
} tm_j2_state;

#define TM_J2_ENERGY_PARTS 1024

/* Energy, residual and tangent at displacement disp [n_points*D] for the
 * history in state (read only).  Each output may be NULL:
 *   energy [1]             sum over elements and points of psi * jxw; needs
 *                          energy_ws: double [n_elem + TM_J2_ENERGY_PARTS],
 *                          which also receives the per-element energies [n_elem]
 *   tangent [nnz*D*D]      BSR(D x D) node-pair CSR order; needs pat and
 *                          tslot_ws [n_elem*B*B*D*D]
 *   residual [n_points*D]  dE/du; needs inc and rslot_ws [n_elem*B*D]
 * form: tables (tab_dev) and n_quad of the element rule (form->form unused).
 * Sums are deterministic (fixed order); where |dev eps_tr| = 0 the tangent is
 * the elastic limit (K - 2mu/3, mu). */
int tm_assemble_j2(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_j2_material* mat,
                   const void* disp, const tm_j2_state* state, void* energy_ws, void* energy,
                   const tm_pattern_view* pat, const tm_incidence_view* inc, void* tslot_ws,
                   void* tangent, void* rslot_ws, void* residual, tm_stream_t stream);

/* Radial-return update of state in place at the converged displacement disp
 * (J2Plasticity.update_state, builtin.py:513-561): where the trial yield
 * value is positive, eps_p += dg n and alpha += sqrt(2/3) dg, with
 * n = dev eps_tr / |dev eps_tr| (a norm below 1e-12 divides by 1). */
int tm_j2_update_state(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_j2_material* mat,
                       const void* disp, const tm_j2_state* state, tm_stream_t stream);

/* ---------------------------------------------------------------------
 * Penalty contact against a plane on the selected boundary facets.
 * Replaces ContactAssembler.energy (tensormesh/assemble/builtin.py:609-728)
 * with the penalty density of its example (builtin.py:588-597) for the plane
 * n . x = offset, and the autograd derivatives of that energy:
 *   x_q = sum_b phi_b(q) x_b, g_q = n . x_q - offset,
 *   energy = sum_facets sum_q kappa/2 max(-g_q, 0)^2 FxW_q,
 * with FxW of the configuration mesh->points and x the positions.
 * ------------------------------------------------------------------- */
typedef struct tm_contact_plane {
  double kappa;              /* penalty stiffness */
  double normal[3];          /* unit obstacle normal n (entries past D unused) */
  double offset;             /* g = n . x - offset; contact where g <= 0 */
} tm_contact_plane;

/* x [n_points*D]: positions in the mesh dtype.  Each output may be NULL:
 *   energy [1]           needs energy_ws: double [n_sel + 1024] (per-facet
 *                        energies, then partial sums)
 *   grad [n_points*D]    dE/dx; needs inc (the selected facets' cell nodes,
 *                        entries f*B + b) and rslot_ws [n_sel*B*D]
 *   hess [nnz*D*D]       d2E/dx2 in the BSR(D x D) node-pair order of the
 *                        cell mesh's pattern; fpat->seg_ptr [nnz + 1] and
 *                        fpat->seg_slot list, per pattern entry, the facet
 *                        slots f*B*B + b*B + c that land on it (ascending);
 *                        entries with none come out 0.  With fpat->entry and
 *                        entry_t (= entry) set, fpat lists only the entries
 *                        that facets touch and only those are written: the
 *                        caller zeroes hess.  tslot_ws [n_sel*B*B*D*D]
 * facet_tab_dev / n_quad as tm_assemble_facet_linear.  With no selected facet
 * (n_sel = 0) energy and grad are set to zero, and so is hess for a full view
 * (fpat->entry NULL); with a subset view hess is left to the caller's zeroing,
 * as in every other call. */
int tm_contact_penalty(const tm_mesh_view* mesh, const tm_facet_view* facets, const double* facet_tab_dev,
                       int32_t n_quad, const tm_contact_plane* plane, const void* x, void* energy_ws,
                       void* energy, const tm_incidence_view* inc, void* rslot_ws, void* grad,
                       const tm_pattern_view* fpat, void* tslot_ws, void* hess, tm_stream_t stream);

/* Row-centric stage S for P1 (TRI3 / TET4): the entries of tm_assemble_bilinear
 * from a stage-G slot buffer, each CSR row reading its incidences' local rows
 * (B contiguous values each) and routing them by colpos; same per-entry order
 * (ascending element).  inc_order = 0: slot in element order (tm_element_matrices);
 * 1: slot in incidence order (tm_element_matrices_inc).  Replaces ReduceProjector.__call__ (projector.py:95-123). */
int tm_reduce_p1_rows(int32_t dtype, int32_t cell, const tm_rows_view* rows, const void* slot, int32_t inc_order,
                      void* values, tm_stream_t stream);

/* Two-stage P1 assembly in incidence order (TRI3 / TET4 scalar forms): stage G
 * writes the local row K_e[i][.] to slot_ws[inc_inv[e*B+i] * B] (inc_inv: the
 * inverse of the incidence permutation, int32 [n_elem*B]); stage S reads each
 * CSR row's contributions contiguously and routes them by colpos / entry masks.
 * Same values and per-entry order as tm_assemble_bilinear.  slot_ws: [n_elem*B*B]. */
int tm_assemble_p1_incidence(const tm_mesh_view* mesh, const tm_form_desc* form, const tm_rows_view* rows,
                             const int32_t* inc_inv, void* slot_ws, void* values, tm_stream_t stream);
/* Its stage G alone. */
int tm_element_matrices_inc(const tm_mesh_view* mesh, const tm_form_desc* form, const int32_t* inc_inv,
                            void* slot_ws, tm_stream_t stream);

/* Stage G alone (element matrices into slot_ws).  Same citation as above. */
int tm_element_matrices(const tm_mesh_view* mesh, const tm_form_desc* form, void* slot_ws,
                        tm_stream_t stream);

/* Stage S alone: out[k*HH + t] = sum over s in seg(k) of slot[s*HH + t],
 * added left to right in segment order.  The deterministic form of
 * ReduceProjector.__call__ (projector.py:95-123).  With a plan (n_blocks > 0)
 * each block gathers its contributions into LDS first; without one, one
 * thread walks each entry. */
int tm_segmented_reduce(int32_t dtype, int64_t nnz, int32_t hh, const int32_t* seg_ptr,
                        const int32_t* seg_slot, int64_t n_blocks, const int32_t* blk_ptr,
                        const void* slot, void* out, tm_stream_t stream);

/* Segment list for symmetric element blocks (TM_FLAG_UPPER), built once per
 * pattern: the slot s = e*B*B + i*B + j of seg_slot[t] becomes the packed
 * upper slot e*B(B+1)/2 + p(a, b) of (a, b) = (min(i, j), max(i, j)),
 * p(a, b) = a*B - a(a-1)/2 + b - a, with bit 31 (0x80000000) set when i > j:
 * the transposed block, which tm_segmented_reduce reads as value b*D + a for
 * value a*D + b of a D x D entry (hh = 4 or 9; bit 31 only selects the slot
 * for other hh).
 * Segment order, and so the order of every sum, is unchanged.  No
 * counterpart in the reference, which stores all B*B blocks
 * (projector.py:95-123 sums them).  n_slots = n_elem * n_basis^2. */
int tm_symmetric_segments(int32_t n_basis, int64_t n_slots, const int32_t* seg_slot, int32_t* sym_slot,
                          tm_stream_t stream);

/* Slot positions for contiguous contributions (tm_pattern_view.slot_dest)
 * from a TM_FLAG_UPPER segment list of n contributions (each packed upper
 * block listed once, as in the paired list of the upper entries):
 * slot_dest[seg_slot[t] & INT32_MAX] = t | (seg_slot[t] & INT32_MIN).
 * Needs elements without repeated vertices (every packed block then
 * appears exactly once).  Device pointers.  The reference has no
 * counterpart: it is the inverse of the ReduceProjector permutation
 * (projector.py:95-123) for the symmetric storage of this build. */
int tm_slot_dest_build(int64_t n, const int32_t* seg_slot, int32_t* slot_dest, tm_stream_t stream);

/* Stage S with paired outputs: segment k (contributions seg_slot[seg_ptr[k]
 * .. seg_ptr[k+1]), summed left to right as tm_segmented_reduce does) is
 * written to out[entry[k]*hh + t] and, when entry_t[k] != entry[k], to
 * out[entry_t[k]*hh + t'] with t' the transposed position of a D x D block
 * (t for hh other than 4 and 9).  With the upper entries (row <= column) of
 * a symmetric matrix and their mirrors, every slot block is read once.  No
 * counterpart in the reference (projector.py:95-123 sums every entry). */
int tm_segmented_reduce_pairs(int32_t dtype, int64_t n_seg, int32_t hh, const int32_t* seg_ptr,
                              const int32_t* seg_slot, const int32_t* entry, const int32_t* entry_t,
                              int64_t n_blocks, const int32_t* blk_ptr, const void* slot, void* out,
                              tm_stream_t stream);

/* ---------------------------------------------------------------------
 * Linear-form assembly (load vector), two stages on one stream.
 * Replaces NodeAssembler.__call__ (tensormesh/assemble/node_assembler.py:357-409)
 * for forward(v, f) = f * v with nodal f (form->coeff, [n_points]) or
 * forward(v) = c * v (const_value != NULL, builtin.py:730-770):
 *   b[n] = sum_{(e,i): conn[e,i]=n} sum_q w_q |det J_q| phi_i(q) f_h(q)
 * slot_ws: [n_elem*B]; out: [n_points]; form->tab_dev and n_quad required.
 * ------------------------------------------------------------------- */
int tm_assemble_linear(const tm_mesh_view* mesh, const tm_form_desc* form, const double* const_value,
                       const tm_incidence_view* inc, void* slot_ws, void* out, tm_stream_t stream);

/* Allocation-free variant of tm_assemble_linear.  No counterpart in the
 * reference: it exists for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and summation order as
 * tm_assemble_linear without slot_ws: the n_elem*B slot values live in a
 * library-owned device array of TM_LINEAR_SCRATCH_BYTES, one per GPU the
 * library runs on, held until the process ends (its address is looked up on
 * the first call on each device).  Calls whose slot values exceed it
 * (n_elem*B*sizeof(value) > TM_LINEAR_SCRATCH_BYTES, e.g. more than 2,097,152
 * float64 tetrahedra) fail with TM_ERR_SIZE before any launch.  The stream
 * must belong to the current device.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (two streams, or two host threads)
 * race; calls in one stream are ordered and safe.  Nothing needs clearing:
 * each call writes every slot value it reads. */
#define TM_LINEAR_SCRATCH_BYTES (64LL << 20)
int tm_assemble_linear_noalloc(const tm_mesh_view* mesh, const tm_form_desc* form, const double* const_value,
                               const tm_incidence_view* inc, void* out, tm_stream_t stream);

/* ---------------------------------------------------------------------
 * Condensation values.  Replaces the per-call part of Condenser.__call__
 * (tensormesh/operator/condense.py:278-296).  Topology arrays come from the
 * once-per-pattern layout (condense.py:182-220), all int32, device:
 *   inner_rp[n_inner+1], inner_idx[nnz_ii]   entries of K_ii per inner row
 *   ou_rp[n_inner+1], ou_idx[nnz_io], ou_col[nnz_io]  entries / outer dof of K_io
 *   inner_dof[n_inner]                        global dof of each inner dof
 * n_kept = inner_rp[n_inner] (the number of K_ii entries, passed so that the
 * call never reads the device).
 * Outputs kii[n_kept] = values[inner_idx], f_i[n_inner] = rhs[inner_dof] - K_io u_o.
 * ------------------------------------------------------------------- */
int tm_condense(int32_t dtype, int64_t n_inner, int64_t n_kept, const void* values, const int32_t* inner_rp,
                const int32_t* inner_idx, const int32_t* ou_rp, const int32_t* ou_idx, const int32_t* ou_col,
                const int32_t* inner_dof, const void* rhs, const void* u_o, void* kii, void* f_i,
                tm_stream_t stream);

/* Boundary-facet load vector (FacetAssembler.__call__, facet_assembler.py:123-270)
 * for the integrands f * v (data: nodal f [n_points * n_comp]) or c * v
 * (cval_dev: n_comp device doubles; data NULL).  facet_tab_dev: float64
 * w [Qf] | phi [nF][Qf][B] | dphi [nF][Qf][D][B] | dgamma [nF][D-1][D] (the
 * cell basis at the facet points mapped into the reference cell, and the
 * reference tangents of each facet).  inc: the facet-node incidence lists of
 * the selected facets' cell nodes (tm_incidence_build on elem's rows, entry
 * facet*B + b), slot_ws [n_sel*B*n_comp], out [n_points*n_comp].  n_comp <= 4.
 * With no selected facet (n_sel = 0) out is set to zero. */
int tm_assemble_facet_linear(const tm_mesh_view* mesh, const tm_facet_view* facets, const double* facet_tab_dev,
                             int32_t n_quad, int32_t n_comp, const void* data, const double* cval_dev,
                             const tm_incidence_view* inc, void* slot_ws, void* out, tm_stream_t stream);

/* ---------------------------------------------------------------------
 * Sparse solve.  The reference inherits solve / @ from torch_sla (not
 * vendored); its in-tree fallback is the Jacobi-CG of
 * tensormesh/sparse/solve/torch_solve.py:54-128.
 *   tm_spmv:        y = A x (CSR; every row summed in ascending entry order)
 *   tm_jacobi:      dinv[i] = 1 / A_ii (1 where A_ii == 0)
 *   tm_pcg_init:    r = b - A x, z = dinv r, p = z, per-system scalars
 *   tm_pcg_iterate: n_iter PCG steps for every still-active system; no host
 *                   synchronisation, so the caller may capture it in a graph
 *                   and poll view->scalars between chunks.  With two_launch
 *                   set, each step is two launches (Chronopoulos-Gear
 *                   recurrences; every workgroup re-sums its system's
 *                   partials in workgroup order, so a step costs
 *                   O(workgroups per system) extra loads per workgroup:
 *                   meant for systems of up to ~2048 workgroups); with
 *                   two_launch = 2 one launch per step (the same
 *                   recurrences: each workgroup forms the next u = M r of
 *                   its SpMV columns itself from the previous step's r,
 *                   A u and A p, double-buffered by parity); otherwise five.  Host-visible scalars per system s (stride 8 for
 *                   five launches, 16 for two): [4] |r|^2, [5] |b|^2,
 *                   [6] active, [7] iterations.
 * ------------------------------------------------------------------- */
int tm_spmv(int32_t dtype, int64_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
            const void* x, void* y, tm_stream_t stream);
int tm_jacobi(int32_t dtype, int64_t n_rows, const int32_t* row_ptr, const int32_t* col_idx, const void* values,
              void* dinv, tm_stream_t stream);
int tm_pcg_init(const tm_pcg_view* view, const void* b, const void* x, tm_stream_t stream);
int tm_pcg_iterate(const tm_pcg_view* view, void* x, int32_t n_iter, tm_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TENSORMESH_AMD_H */
