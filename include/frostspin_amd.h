/*
 * frostspin_amd C-ABI: block-sparse SymmetricTensor engine for MI355X (gfx950).
 *
 * Every entry point returns an int status (FS_OK on success); on failure the
 * message is available from fs_last_error() (thread-local). No C++ exception
 * crosses this boundary. Device pointers are plain `double*` into memory the
 * caller owns; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * The library never allocates or frees caller tensors. Plans own the device
 * tables they upload and release them in fs_permute_plan_destroy.
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line relative to the frostspin source tree.
 */
#ifndef FROSTSPIN_AMD_H
#define FROSTSPIN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_OK 0
#define FS_ERR_INVALID 1
#define FS_ERR_HIP 2
#define FS_ERR_NOMEM 3

/* ------------------------------------------------------------------------- */
/* Library                                                                   */
/* ------------------------------------------------------------------------- */

/* Last error message of the calling thread ("" if none). */
const char *fs_last_error(void);

/* ABI version of this header (bumped on any signature change). */
int fs_abi_version(void);

/* Number of visible HIP devices (0 with no GPU; never fails on a CPU host). */
int fs_device_count(int *count);

/* ------------------------------------------------------------------------- */
/* Host planner (no GPU needed)                                              */
/* ------------------------------------------------------------------------- */

/*
 * U(1) charge fusion: out = Kronecker sum over legs of (flip[l] ? -c_l : c_l),
 * row-major (last leg fastest), with int8 wrap-around (Z2 as {0,-128}).
 * out has prod(sizes) entries.
 * Replaces _numba_combine_U1 / U1SymmetricTensor.combine_representations
 * (symmetric_tensor/u1_symmetric_tensor.py:8-19, 35-40).
 */
int fs_u1_fuse(int nlegs, const int8_t *const *legs, const int64_t *sizes,
               const uint8_t *flip, int8_t *out);

/*
 * Greedy multiplet-aware truncation over blockwise non-increasing values.
 * values[b] has sizes[b] entries; dims[b] is the irrep dimension counted per
 * kept value; cutoff = atol + rtol * max value (already combined by caller).
 * cuts[b] receives the number of values kept in block b.
 * Replaces _numba_find_block_cuts (misc_tools/svd_tools.py:57-74).
 */
int fs_find_block_cuts(int nblocks, const double *const *values,
                       const int64_t *sizes, int64_t cut_target,
                       const int64_t *dims, double cutoff, double degen_ratio,
                       int64_t *cuts);

/* ------------------------------------------------------------------------- */
/* Abelian permute plan                                                      */
/* ------------------------------------------------------------------------- */

typedef struct fs_permute_plan fs_permute_plan;

/*
 * Build the index maps of an abelian (U(1), Z2-as-int8 or trivial) permute.
 *   ndim, shape[ndim]       old tensor shape (leg dimensions)
 *   leg_charges[ndim]       int8 charges of each old leg (shape[l] entries)
 *   signature[ndim]         old signature (1 = flipped charge)
 *   old_nrr                 number of old row legs
 *   old_nblocks, old_block_irreps   blocks present in the old tensor
 *                           (sorted; missing allowed blocks read as zero)
 *   old_block_offsets       start of each old block in the source buffer
 *                           (NULL: packed in block order)
 *   old_layout_f            1 if old blocks are stored column-major
 *   old_dense               1: old data is one dense row-major matrix
 *                           (prod row dims x prod col dims); charges only
 *                           shape the new side (from_array)
 *   axes[ndim], new_nrr     permutation (new leg i = old leg axes[i])
 *   new_dense               1: write one dense row-major matrix (toarray)
 * Replaces AbelianSymmetricTensor._permute_data + _numba_abelian_transpose
 * (symmetric_tensor/abelian_symmetric_tensor.py:395-430, 103-250),
 * _blocks_from_dense/_numba_reduce_to_blocks (:363-373, 19-40) and
 * _tomatrix/_numba_blocks_to_array (:375-383, 63-75).
 */
int fs_abelian_permute_plan_create(int ndim, const int64_t *shape,
                                   const int8_t *const *leg_charges,
                                   const uint8_t *signature, int old_nrr,
                                   int old_nblocks,
                                   const int8_t *old_block_irreps,
                                   const int64_t *old_block_offsets,
                                   int old_layout_f, int old_dense,
                                   const int32_t *axes, int new_nrr,
                                   int new_dense, fs_permute_plan **plan);

/* Number of output blocks and stored coefficients. */
int fs_permute_plan_info(const fs_permute_plan *plan, int64_t *new_nblocks,
                         int64_t *new_ncoeff, int64_t *old_ncoeff);

/* Output block irreps (sorted) and shapes (nblocks x 2, row-major). */
int fs_permute_plan_blocks(const fs_permute_plan *plan, int8_t *irreps,
                           int64_t *shapes);

/*
 * The plan is a list of boxes. A box holds the elements whose index on every
 * leg lies in one run of equal charges of that leg; inside it the old and new
 * in-block positions are affine in the legs' digits, so a tile (a box or a
 * slice of one along its slowest destination digits, at most 4096 elements)
 * is a dense strided copy staged through LDS. The plan stores one record per
 * box; the kernel derives each tile from its box. Host copy of the tiles as
 * the kernel derives them (test/inspection): *ntiles and *nboxes always; if
 * out is not NULL it receives ntiles records of FS_PERMUTE_TILE_RECORD int64
 * each:
 *   src_base (-1: absent old block, zeros), dst_base, vol, ns, nd,
 *   ns x (len, src stride, staging stride) padded to FS_PERMUTE_MAX_DIGITS,
 *   nd x (len, dst stride) padded to FS_PERMUTE_MAX_DIGITS;
 * source digits fastest first (flat tile index f in source order: digit k is
 * (f / prod of the faster lengths) mod len), staging offsets in destination
 * order, destination digits fastest first likewise.
 */
#define FS_PERMUTE_MAX_DIGITS 12
#define FS_PERMUTE_TILE_RECORD (5 + 5 * FS_PERMUTE_MAX_DIGITS)
int fs_permute_plan_tiles(const fs_permute_plan *plan, int64_t *ntiles, int64_t *nboxes,
                          int64_t *out);

/* Gather old blocks (packed, in block order) into the new packed blocks. */
int fs_permute_plan_execute_f64(fs_permute_plan *plan, const double *src,
                                double *dst, void *stream);

/* Same map, FP32 storage (mixed-precision form). */
int fs_permute_plan_execute_f32(fs_permute_plan *plan, const float *src,
                                float *dst, void *stream);

int fs_permute_plan_destroy(fs_permute_plan *plan);

/* ------------------------------------------------------------------------- */
/* Structural transforms of non-abelian (SU(2)) tensors                      */
/* ------------------------------------------------------------------------- */

/*
 * Data movement of Lie-group tensors (SU(2)) in one planned kernel. A plan is
 * a list of tasks; a task is one elementary ("simple") block: one irrep per
 * leg, E = product of the legs' external degeneracies. Its coefficients form
 * n_in input rows and n_out output rows of E entries each (structural x
 * external), and the task computes
 *
 *     out[o, e] = sum_i W[o, i] * in[i, e],      e = 0 .. E-1,
 *
 * where e is a row-major multi-index over the task's digits (one per leg,
 * radix = that leg's external degeneracy). Digit k contributes
 * digit*in_r / digit*in_c to the input row/column index and digit*out_r /
 * digit*out_c to the output ones; the element of row entry r at (er, ec) is
 * at r.base + er*r.rs + ec*r.cs. An input row with base < 0 is an absent
 * block and reads as zero. W is stored transposed per task: W^T[i, o] at
 * w[w_off + i*n_out_pad + o], n_out_pad a multiple of 8 >= n_out.
 *
 * Permute (W = CG-tree overlap unitaries, lie_group_symmetric_tensor.py:
 * 96-265, 432-835, 1116-1134), dense -> blocks (W = CG projector / irrep
 * dimension, lie:856-967) and blocks -> dense (W = CG projector, lie:969-1109)
 * are all plans of this form; every output coefficient is written exactly
 * once by the permute plans, so their destination needs no initialisation.
 *
 * fs_lie_plan_create validates every access against src_size / dst_size
 * (in doubles), copies the tables to the current device and returns an
 * opaque plan.
 */
#define FS_LIE_MAX_LEGS 16

typedef struct fs_lie_task {
  int64_t e_size;
  int64_t w_off;
  int64_t n_in;
  int64_t n_out;
  int64_t n_out_pad;
  int64_t row0;   /* n_in input row entries, then n_out output row entries */
  int64_t ndig;
  int64_t dig0;
} fs_lie_task;

typedef struct fs_lie_row {
  int64_t base;
  int64_t rs;
  int64_t cs;
} fs_lie_row;

typedef struct fs_lie_digit {
  int64_t radix;
  int64_t in_r;
  int64_t in_c;
  int64_t out_r;
  int64_t out_c;
} fs_lie_digit;

typedef struct fs_lie_plan fs_lie_plan;

int fs_lie_plan_create(int64_t ntasks, const fs_lie_task *tasks, int64_t nrows,
                       const fs_lie_row *rows, int64_t ndigits, const fs_lie_digit *digits,
                       int64_t nw, const double *w, int64_t src_size, int64_t dst_size,
                       fs_lie_plan **plan);

/* Validation of create's tables alone (no device work): FS_OK or
 * FS_ERR_INVALID with the reason in fs_last_error(). */
int fs_lie_plan_check(int64_t ntasks, const fs_lie_task *tasks, int64_t nrows,
                      const fs_lie_row *rows, int64_t ndigits, const fs_lie_digit *digits,
                      int64_t nw, int64_t src_size, int64_t dst_size);

/* Tables of a structural-transform plan for one permute, from the irreps-only
 * pattern of the permute and the degeneracies of the legs. Replaces the
 * per-call offset computation of the reference's numba permute
 * (lie_group_symmetric_tensor.py:96-265: the in-block row / column of every
 * elementary block and internal index, old and new side) by one host pass.
 *
 * The pattern (built once per irreps pattern; every array int64, row-major):
 *   ndim, nrr, nrr_out      legs, row legs before and after the permute
 *   ntasks                  elementary blocks (one irrep per leg) with data
 *   axes[ndim]              new leg k is old leg axes[k]
 *   nirr[ndim]              irreps per old leg
 *   sig_idx[ntasks][ndim]   irrep index of each old leg of a task
 *   ir_in, ic_in [ntasks]   row / column irreps combination (row-major over the
 *                           old row / column legs); ir_out, ic_out on the new side
 *   n_in, n_out, w_off, n_out_pad [ntasks]   as in fs_lie_task
 *   nt_in, nt_out           target blocks of the old / new side
 *   idr_in[rc_in][nt_in], idc_in[cc_in][nt_in]   internal degeneracy of each
 *                           row / column combination in each target block
 *                           (rc_in = prod of nirr over the old row legs, ...);
 *                           idr_out, idc_out on the new side
 *   in_t, in_a, in_b [sum n_in]      per input row of every task, in task order:
 *                           target block, internal row, internal column;
 *   out_t, out_a, out_b [sum n_out]  the same for the output rows
 * degen: the degeneracies of the old legs, concatenated (nirr[k] each).
 * in_block[nt_in]: stored block holding each old target (-1: absent, its rows
 * read as zero); block_off / block_rs / block_cs [nblocks]: offset and strides
 * of the stored blocks. The new side is packed row-major in target order.
 * Outputs: tasks [ntasks][8], rows [sum(n_in + n_out)][3],
 * digits [ntasks * ndim][5] (fs_lie_task / fs_lie_row / fs_lie_digit layouts)
 * and the new block shapes [nt_out][2]. ndegen and nrows are the lengths of
 * degen and of rows (in entries); the other arrays are sized from the pattern
 * and every index is checked against it. */
typedef struct fs_lie_pattern {
  int64_t ndim, nrr, nrr_out, ntasks, nt_in, nt_out;
  const int64_t *axes, *nirr, *sig_idx;
  const int64_t *ir_in, *ic_in, *ir_out, *ic_out;
  const int64_t *n_in, *n_out, *w_off, *n_out_pad;
  const int64_t *idr_in, *idc_in, *idr_out, *idc_out;
  const int64_t *in_t, *in_a, *in_b, *out_t, *out_a, *out_b;
} fs_lie_pattern;

int fs_lie_plan_tables(const fs_lie_pattern *pattern, int64_t ndegen, const int64_t *degen,
                       const int64_t *in_block, int64_t nblocks, const int64_t *block_off,
                       const int64_t *block_rs, const int64_t *block_cs, int64_t *tasks,
                       int64_t nrows, int64_t *rows, int64_t *digits, int64_t *shapes);

/* Structural matrices of one pattern (W as for fs_lie_plan_create, indexed by
 * the pattern's w_off / n_in / n_out_pad) uploaded once to the current device
 * and shared by every plan built from that pattern: they depend on the irreps
 * only, not on the degeneracies. Must outlive the plans that use it. */
typedef struct fs_lie_weights fs_lie_weights;

int fs_lie_weights_create(const fs_lie_pattern *pattern, int64_t nw, const double *w,
                          fs_lie_weights **weights);
int fs_lie_weights_destroy(fs_lie_weights *weights);

/* fs_lie_plan_tables followed by plan creation in one call: the plan of one
 * permute straight from its pattern and degeneracies, the tables staying on
 * the host side of the library and W taken from the pattern's weights.
 * shapes: the new block shapes [nt_out][2]; the destination size is the sum
 * of their products. */
int fs_lie_plan_build(const fs_lie_pattern *pattern, int64_t ndegen, const int64_t *degen,
                      const int64_t *in_block, int64_t nblocks, const int64_t *block_off,
                      const int64_t *block_rs, const int64_t *block_cs,
                      const fs_lie_weights *weights, int64_t src_size, int64_t *shapes,
                      fs_lie_plan **plan);

/* Work of one execution: flops (2 * sum n_in*n_out*E) and coefficients read
 * and written (sum (n_in + n_out) * E). */
int fs_lie_plan_info(const fs_lie_plan *plan, int64_t *ntiles, double *flops, double *coeffs);

int fs_lie_plan_execute_f64(const fs_lie_plan *plan, const double *src, double *dst,
                            void *stream);

int fs_lie_plan_destroy(fs_lie_plan *plan);

/* ------------------------------------------------------------------------- */
/* Grouped GEMM over uneven blocks                                           */
/* ------------------------------------------------------------------------- */

/*
 * One group: C[m x n] = alpha * A[m x k] * B[k x n] + beta * C.
 * Element (i,j) of A is at A + a_off + i*a_rs + j*a_cs (same for B);
 * C is row-major with leading dimension c_ld (C + c_off + i*c_ld + j).
 * Replaces the per-block numpy matmul loop of SymmetricTensor.__matmul__
 * (symmetric_tensor/symmetric_tensor.py:453-489) and the sector products of
 * construct_projectors (ctmrg/ctm_renormalize.py:92-94, 122-123).
 */
typedef struct fs_gemm_group {
  int64_t m, n, k;
  int64_t a_off, a_rs, a_cs;
  int64_t b_off, b_rs, b_cs;
  int64_t c_off, c_ld;
} fs_gemm_group;

/*
 * FP64 MFMA (v_mfma_f64_16x16x4f64), FP64 accumulation. Groups with at most
 * 16 rows, K <= 128 and n >= 256 run on a streaming (HBM-bound) kernel;
 * launches whose tiles cannot fill the chip while K is long are split over K
 * (partials summed in a fixed order, so results are reproducible).
 */
int fs_grouped_gemm_f64(int ngroups, const fs_gemm_group *groups,
                        const double *A, const double *B, double *C,
                        double alpha, double beta, void *stream);

/*
 * Same product with per-group operand bases: group i reads A[i] + a_off,
 * B[i] + b_off and writes C[i] + c_off (host arrays of ngroups device
 * pointers). One launch then covers blocks that live in different tensors,
 * e.g. the sector products of several independent projector computations
 * (construct_projectors, ctm_renormalize.py:6-136, called once per site and
 * direction by the CTMRG schedules).
 */
int fs_grouped_gemm_f64_multi(int ngroups, const fs_gemm_group *groups,
                              const double *const *A, const double *const *B,
                              double *const *C, double alpha, double beta,
                              void *stream);

/*
 * fs_grouped_gemm_f64 without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same kernels and
 * results as fs_grouped_gemm_f64; the split-K partial tiles go to a static
 * per-device scratch array of 2^21 doubles (16 MiB, held by each device the
 * library runs on until the process ends) instead of a pooled buffer. A call
 * whose split-K partials would exceed it fails with FS_ERR_INVALID (the
 * split rule keeps them within the resident workgroup slots plus the output
 * tiles: up to ~18 MiB with 32 x 32 tiles, more with 64 x 64 tiles over a
 * long K, which are then refused). Every partial is written before
 * it is read, so nothing needs clearing between calls. Calls that overlap on
 * one device (other streams, other threads' streams) share the array and
 * race; serialize them.
 */
int fs_grouped_gemm_f64_noalloc(int ngroups, const fs_gemm_group *groups,
                                const double *A, const double *B, double *C,
                                double alpha, double beta, void *stream);

/*
 * Kernel dispatches issued by the calling thread's last grouped GEMM call
 * (tile kernel per 32 groups, skinny-row kernel, split-K reduction); for
 * per-dispatch timing. No counterpart in the reference (instrumentation).
 */
int fs_gemm_last_dispatches(void);

/*
 * Compensated FP32 form: A and B are split into hi/lo FP32 parts on the fly
 * (x = hi + lo), the three products hi*hi + hi*lo + lo*hi run on FP32 MFMA
 * with FP32 accumulation per K chunk and FP64 accumulation across chunks.
 * Inputs and outputs stay FP64.
 */
int fs_grouped_gemm_f64_split(int ngroups, const fs_gemm_group *groups,
                              const double *A, const double *B, double *C,
                              double alpha, double beta, void *stream);

/* ------------------------------------------------------------------------- */
/* Batched FP64 SVD (one-sided Jacobi)                                       */
/* ------------------------------------------------------------------------- */

/*
 * A_b (m x n, element (i,j) at A + a_off + i*a_rs + j*a_cs) = U S V^T with
 * q = min(m,n): U (m x q) row-major at U + u_off, S (q) at S + s_off sorted
 * non-increasing, VT (q x n) row-major at VT + vt_off. work must hold
 * work_off .. work_off + q*(max(m,n)+q) doubles per matrix.
 * Replaces robust_svd (misc_tools/svd_tools.py:116-151) per block in
 * SymmetricTensor.svd/truncated_svd (symmetric_tensor.py:691-710, 835-877)
 * and construct_projectors (ctmrg/ctm_renormalize.py:96-103).
 */
typedef struct fs_svd_desc {
  int64_t m, n;
  int64_t a_off, a_rs, a_cs;
  int64_t u_off, s_off, vt_off, work_off;
} fs_svd_desc;

int fs_batched_svd_f64(int nmat, const fs_svd_desc *descs, const double *A,
                       double *U, double *S, double *VT, double *work,
                       int max_sweeps, void *stream);

/* ------------------------------------------------------------------------- */
/* Batched FP64 Householder QR                                               */
/* ------------------------------------------------------------------------- */

/*
 * A_b (m x n, strided as in fs_svd_desc) = Q R with q = min(m,n): Q (m x q)
 * row-major at Q + q_off, R (q x n) row-major upper triangular at R + r_off
 * (R may be NULL). LAPACK geqrf/orgqr sign conventions. work must hold
 * m*n + q + m*q doubles per matrix from work_off.
 * Replaces the per-block scipy.linalg.qr(mode="economic") of
 * SymmetricTensor.qr (symmetric_tensor.py:655-671).
 */
typedef struct fs_qr_desc {
  int64_t m, n;
  int64_t a_off, a_rs, a_cs;
  int64_t q_off, r_off, work_off;
} fs_qr_desc;

int fs_batched_qr_f64(int nmat, const fs_qr_desc *descs, const double *A,
                      double *Q, double *R, double *work, void *stream);

/* ------------------------------------------------------------------------- */
/* CholeskyQR building blocks (randomized truncated SVD)                     */
/* ------------------------------------------------------------------------- */

/*
 * Upper Cholesky G + s I = R^T R of symmetric positive (semi)definite
 * l x l matrices (row-major at G + g_off, upper triangle read), with
 * s = shift_rel * trace(G); R written row-major at R + r_off. status[b] (device
 * int32) = 0 on success, j+1 if pivot j was not positive. These replace the
 * orthonormalisation that ARPACK performs inside sparse_svd
 * (misc_tools/svd_tools.py:154-197) for the randomized range finder.
 */
typedef struct fs_chol_desc {
  int64_t l, g_off, r_off;
  double shift_rel;
} fs_chol_desc;

int fs_batched_cholesky_f64(int nmat, const fs_chol_desc *descs,
                            const double *G, double *R, int32_t *status,
                            void *stream);

/*
 * O = Y R^-1 with R upper triangular (l x l row-major at R + r_off), Y (m x l)
 * strided (y_rs, y_cs) at Y + y_off, O row-major (m x l) at O + o_off.
 */
/*
 * Cholesky and triangular inverse in one launch: R = chol(sym(G) + shift I)
 * (upper, G = R^T R) and Rinv = R^-1, both row-major l x l, one workgroup per
 * matrix in LDS (l <= 128). R may be NULL (not written). status[i] = 0 on
 * success, j+1 when pivot j is not positive (R, Rinv unspecified). The solve
 * Y R^-1 of CholeskyQR then runs as a grouped GEMM with Rinv. Replaces the
 * QR step of the reference's randomized/ARPACK sector SVDs like
 * fs_batched_cholesky_f64 (misc_tools/svd_tools.py:154-197).
 * R or Rinv may be NULL (not both): with Rinv NULL only the factor R is
 * written (the factorisation alone; fs_batched_triinv_f64 then inverts it).
 */
typedef struct fs_cholinv_desc {
  int64_t l;
  int64_t g_off;     /* G = sum_{p < n_part} of the l x l matrices at g_off + p*g_stride */
  int64_t g_stride;  /* (partial Gram products of a split-K GEMM) */
  int64_t n_part;
  int64_t r_off, rinv_off;
  double shift_rel;
} fs_cholinv_desc;

int fs_batched_cholinv_f64(int nmat, const fs_cholinv_desc *descs, const double *G,
                           double *R, double *Rinv, int32_t *status, void *stream);

/*
 * Inverse X = R^-1 of upper triangular matrices (row-major l x l at R + r_off,
 * X row-major at X + x_off, l <= 128), one workgroup per matrix, blocked by 16
 * rows. status (may be NULL) is the per-matrix status of the factorisation
 * that produced R: a nonzero entry makes that X all NaN. With the factor-only
 * form of fs_batched_cholinv_f64 this gives R and R^-1 for CholeskyQR
 * (misc_tools/svd_tools.py:154-197 of the reference orthonormalises with
 * Householder QR; no counterpart of this split there).
 */
typedef struct fs_triinv_desc {
  int64_t l;
  int64_t r_off, x_off;
} fs_triinv_desc;

int fs_batched_triinv_f64(int nmat, const fs_triinv_desc *descs, const double *R, double *X,
                          const int32_t *status, void *stream);

/*
 * Multi-pass shifted CholeskyQR of tall matrices in one call: Y_i (m x l,
 * 1 <= l <= 128, element (r, c) at Y + y_off + r*y_rs + c*y_cs) -> Q_i (m x l
 * row-major at Q + q_off) with orthonormal columns. Each pass issues, on one
 * stream, the split-K Gram product G = Y^T Y (fs_grouped_gemm_f64; row chunks
 * of 192, at most 16), the factor-only fs_batched_cholinv_f64 of G + s I
 * (s = 11 (m l + l (l + 1)) eps relative on the first pass when shifted),
 * fs_batched_triinv_f64 and Y <- Y R^-1 (fs_grouped_gemm_f64); the next pass
 * reads the packed result. Rt (may be NULL; needs passes >= 2) receives
 * R = R_{passes-1} ... R_0 (l x l row-major at Rt + rt_off), so Y = Q R.
 * status (passes * nmat int32) holds each pass's factorisation status; a
 * breakdown turns that matrix's Q into NaN, leaves its R unspecified and
 * leaves the fallback to the caller. work holds fs_cholqr_work_size doubles. The orthonormalisations of
 * the randomized SVD (misc_tools/svd_tools.py:154-197 of the reference, which
 * uses Householder QR there); no single counterpart in the reference.
 */
typedef struct fs_cholqr_desc {
  int64_t m, l;
  int64_t y_off, y_rs, y_cs;
  int64_t q_off, rt_off;
} fs_cholqr_desc;

int fs_cholqr_work_size(int nmat, const fs_cholqr_desc *descs, int passes, int want_r,
                        int64_t *doubles);

int fs_cholqr_f64(int nmat, const fs_cholqr_desc *descs, const double *Y, double *Q, double *Rt,
                  double *work, int64_t work_doubles, int passes, int shifted, int32_t *status,
                  void *stream);

typedef struct fs_trsm_desc {
  int64_t m, l, y_off, y_rs, y_cs, r_off, o_off;
} fs_trsm_desc;

int fs_batched_trsm_right_upper_f64(int nmat, const fs_trsm_desc *descs,
                                    const double *Y, const double *R, double *O,
                                    void *stream);

/* ------------------------------------------------------------------------- */
/* Elementwise / reduction helpers on packed block buffers                   */
/* ------------------------------------------------------------------------- */

/*
 * Per-segment weighted sum of squares: out[0] = sum_b w[b] * ||x[off_b:off_b+len_b]||^2
 * computed in a fixed order (deterministic for a given segment list).
 * seg_off (host, nseg+1 entries) and w (host, nseg) are copied into kernel
 * arguments. partial is device scratch of FS_NORM_PARTIALS doubles; out is a
 * device pointer (1 double). Replaces SymmetricTensor.norm
 * (symmetric_tensor.py:624-631).
 */
#define FS_NORM_PARTIALS 1024
int fs_weighted_norm2_f64(int nseg, const int64_t *seg_off, const double *w,
                          const double *x, double *partial, double *out,
                          void *stream);

/* x[i] *= alpha for i < n (in place). Replaces __imul__/__itruediv__
 * (symmetric_tensor.py:441-451). */
int fs_scale_f64(int64_t n, double alpha, double *x, void *stream);

/* x[i] *= 1 / sqrt(d[0]) where d is a device scalar (norm^2): normalises a
 * tensor without a host round trip (nT /= nT.norm(), ctm_renormalize.py:327). */
int fs_scale_rsqrt_dev_f64(int64_t n, const double *d, double *x, void *stream);

/* Marks every plan-table copy that has completed as such (polls the events
 * behind them). No counterpart in the reference: the CTMRG graph recorder
 * calls it after a device synchronisation and before it captures, so that
 * plans executed inside the capture need no event query (events last
 * recorded on a stream that is being captured cannot be queried). */
int fs_plan_tables_settle(void);

/* Stream-ordered copy of `bytes` between device memory and pinned host memory
 * (either direction, hipMemcpyDefault). No counterpart in the reference: the
 * CTMRG graph recorder reads singular values back and writes the projector
 * weights through it, so that those copies become nodes of a recorded graph,
 * at fixed host addresses, instead of going through the framework's host
 * allocator. */
int fs_memcpy_async(void *dst, const void *src, int64_t bytes, void *stream);

/* y[i] = alpha * x[i] + beta * y[i]. Used by __add__/__sub__ (:349-375). */
int fs_axpby_f64(int64_t n, double alpha, const double *x, double beta,
                 double *y, void *stream);

/*
 * Batched strided 2-D copy with optional scaling, column gather and
 * accumulation:
 *   dst[d_off + i*d_rs + j*d_cs] = alpha * src[s_off + i*s_rs + jj*s_cs]
 *                                  * (row_w ? row_w[rw_off+i] : 1)
 *                                  * (col_w ? col_w[cw_off+j] : 1)
 *                                  + beta * dst[...]
 * with jj = col_idx ? col_idx[ci_off+j] : j. row_w, col_w and col_idx are
 * device arrays (may be NULL); beta == 0 never reads dst. Used for
 * transposes, slicing U[:, :k] / V[:k, :], diagonal products (_diagonal_mul,
 * symmetric_tensor.py:1189-1248), block sums (__add__, :349-372), reordering
 * eigenvectors (_sparse_eig, :1316-1322) and projector scaling s^-1/2
 * (ctm_renormalize.py:283-285).
 */
typedef struct fs_copy2d_desc {
  int64_t rows, cols;
  int64_t s_off, s_rs, s_cs;
  int64_t d_off, d_rs, d_cs;
  int64_t rw_off, cw_off, ci_off;
  double alpha, beta;
} fs_copy2d_desc;

int fs_batched_copy2d_f64(int ndesc, const fs_copy2d_desc *descs,
                          const double *src, double *dst, const double *row_w,
                          const double *col_w, const int64_t *col_idx,
                          void *stream);

/*
 * Two-term weighted gather (O(2) data movement): for each descriptor,
 *   out[i][j] = beta*out[i][j] + sum_{a,b<2} w[rw_off+2i+a] w[cw_off+2j+b]
 *               * src[s_off + idx[ri_off+2i+a]*s_rs + idx[ci_off+2j+b]*s_cs]
 * with out[i][j] at dst + d_off + i*d_rs + j*d_cs; source row/column indices
 * lie in [0, src_rows) x [0, src_cols); zero-weight terms are not read.
 * idx and w are device arrays; fs_gather2d_check validates the same
 * descriptors and (host copies of the) index table against the buffer sizes.
 * Replaces split_b0 / _numba_merge_b0oe / _numba_generate_refl_block
 * (o2_symmetric_tensor.py:268-352) and the reflected writes of
 * _numba_O2_transpose (:357-581).
 */
typedef struct fs_gather2d_desc {
  int64_t rows, cols, src_rows, src_cols;
  int64_t s_off, s_rs, s_cs;
  int64_t d_off, d_rs, d_cs;
  int64_t ri_off, rw_off, ci_off, cw_off;
  double beta;
} fs_gather2d_desc;

int fs_gather2d_check(int ndesc, const fs_gather2d_desc *descs, const int64_t *idx, int64_t nidx,
                      int64_t nw, int64_t src_size, int64_t dst_size);

int fs_batched_gather2d_f64(int ndesc, const fs_gather2d_desc *descs, const int64_t *idx,
                            const double *w, const double *src, double *dst, void *stream);

/*
 * Weighted sum of the diagonals: out[0] = sum_b w[b] * trace(block_b) where
 * block b is n_b x n_b at off_b with row stride ld_b (trace, :633-637).
 */
int fs_weighted_trace_f64(int nseg, const int64_t *off, const int64_t *n,
                          const int64_t *ld, const double *w, const double *x,
                          double *out, void *stream);

/*
 * Deterministic counter-based N(0,1) fill (Box-Muller on a 64-bit hash of
 * (seed, index)); used for randomized range finders and synthetic inputs.
 */
int fs_fill_normal_f64(int64_t n, uint64_t seed, double *x, void *stream);

/* FP64 <-> FP32 conversion for the mixed-precision storage form. */
int fs_convert_f64_to_f32(int64_t n, const double *x, float *y, void *stream);
int fs_convert_f32_to_f64(int64_t n, const float *x, double *y, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FROSTSPIN_AMD_H */
