/*
 * qlgpu.h -- C ABI of the MI355X (gfx950) effective-Hamiltonian matvec and
 * Krylov solvers for two-site DMRG / TDVP over U(1) and U(1)xU(1)
 * block-sparse tensors.
 *
 * Conventions
 *   - Every entry point returns an int status (QLG_OK == 0).  Nothing calls
 *     exit() or throws across this boundary; the message of the last failure
 *     on a context is available from qlg_last_error().
 *   - All tensor data and Krylov vectors are device pointers owned by the
 *     caller.  A plan owns its own workspace (intermediates, task lists,
 *     Krylov basis) and frees it in qlg_plan_destroy().
 *   - One HIP stream per context; calls on one context are not thread safe.
 *   - Complex data (QLG_C128) is interleaved (re, im) double pairs.
 *
 * Block-sparse tensor descriptor (qlg_bsten): an index is an ordered list of
 * quantum-number sectors (qn, degeneracy) with a direction; a block is one
 * sector per axis and is stored dense, row-major, at blk_offset[] elements
 * from data.  A block may exist only where sum_OUT qn - sum_IN qn equals the
 * tensor's divergence.  This is the storage model of the reference's
 * QLTensor (TensorToolkit), so a caller can hand its own block map over
 * without copying.
 */
#ifndef QLGPU_H
#define QLGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define QLG_OK 0
#define QLG_ERR_INVALID 1      /* malformed descriptor / argument          */
#define QLG_ERR_HIP 2          /* a HIP runtime call failed                */
#define QLG_ERR_ALLOC 3        /* device allocation failed                 */
#define QLG_ERR_UNSUPPORTED 4  /* valid request this library does not handle */
#define QLG_ERR_NUMERIC 5      /* tridiagonal eigen-solver did not converge*/

/* ---- element / arithmetic types ---------------------------------------- */
#define QLG_F64 0   /* double                                            */
#define QLG_C128 1  /* complex double, interleaved (re, im)              */

/* Arithmetic of the sector GEMMs.  The Krylov vectors, the recurrence, the
 * tridiagonal solve (and the small-matrix exponential) and all reductions
 * stay FP64 in every mode; the reduced modes keep the H_eff intermediates in
 * FP32 and accumulate in FP32.  qlg_lanczos_gs in a reduced mode reports the
 * FP64 Rayleigh quotient of its ground state (one FP64 matvec), so the energy
 * error is second order in the eigenvector error.  Complex plans support
 * QLG_PREC_F64, QLG_PREC_BF16X3 and QLG_PREC_BF16X6 (complex products from
 * real split-piece products: Re.Re - Im.Im and Re.Im + Im.Re). */
#define QLG_PREC_F64 0     /* FP64 MFMA                                          */
#define QLG_PREC_F32 1     /* FP32 operands, FP32 MFMA                           */
#define QLG_PREC_BF16X3 2  /* operands split in 2 bf16 terms (hi + lo), products
                              hi.hi + hi.lo + lo.hi on bf16 MFMA (~16-bit)     */
#define QLG_PREC_BF16X6 3  /* operands split in 3 bf16 terms, the 6 products of
                              index sum <= 2 on bf16 MFMA (~24-bit, FP32-like) */

#define QLG_MAX_RANK 8
#define QLG_MAX_QN_LEN 2

typedef struct qlg_ctx_s *qlg_ctx;
typedef struct qlg_plan_s *qlg_plan;

typedef struct {
  int32_t rank;               /* <= QLG_MAX_RANK                            */
  int32_t qn_len;             /* 1: U(1); 2: U(1)xU(1)                      */
  const int32_t *nsect;       /* [rank] sectors per axis                    */
  const int32_t *sect_qn;     /* [sum(nsect) * qn_len], axis-major          */
  const int64_t *sect_deg;    /* [sum(nsect)] degeneracies                  */
  const int8_t *dir;          /* [rank] -1 = IN, +1 = OUT                   */
  int64_t nblocks;
  const int32_t *blk_coords;  /* [nblocks * rank] sector id per axis        */
  const int64_t *blk_offset;  /* [nblocks] element offset of each block     */
  int32_t dtype;              /* QLG_F64 / QLG_C128                         */
  const void *data;           /* device pointer                             */
} qlg_bsten;

typedef struct {
  double error;               /* LanczosParams::error                      */
  int64_t max_iterations;     /* LanczosParams::max_iterations             */
} qlg_lanczos_params;

typedef struct {
  int64_t iters;              /* LanczosRes::iters                          */
  double gs_eng;              /* LanczosRes::gs_eng                         */
} qlg_lanczos_result;

/* ---- context ------------------------------------------------------------ */
/* stream: a hipStream_t, or NULL for the library's own stream. */
int qlg_ctx_create(int device, void *stream, qlg_ctx *out);
int qlg_ctx_destroy(qlg_ctx ctx);
const char *qlg_last_error(qlg_ctx ctx);
int qlg_ctx_synchronize(qlg_ctx ctx);

/* Block cache of the library's own device buffers (plan workspaces, Krylov
 * bases).  No counterpart in the reference, whose tensors allocate through
 * qlten; here plans are built and destroyed at every site of a sweep, and the
 * cache keeps their buffers for the next plan instead of a hipMalloc / hipFree
 * pair each time.  The cache holds at most min(16 GiB, 1/8 of the device)
 * (QLG_POOL_MB overrides; QLG_POOL=0 at context creation turns it off).
 * qlg_ctx_release_cached returns the cached blocks to the driver (the blocks
 * of live plans stay with them).  qlg_ctx_memory_stats writes
 * {cached bytes, hipMalloc calls, reuses, hipFree calls}; all zero when the
 * cache is off. */
int qlg_ctx_release_cached(qlg_ctx ctx);
int qlg_ctx_memory_stats(qlg_ctx ctx, int64_t stats[4]);

/* ---- device memory plumbing (for hosts without their own HIP allocator) --
 * Allocations are plain hipMalloc memory on the context's device; copies run
 * on the context stream and return after completion.                      */
int qlg_device_alloc(qlg_ctx ctx, int64_t bytes, void **ptr);
int qlg_device_free(qlg_ctx ctx, void *ptr);
int qlg_memcpy_h2d(qlg_ctx ctx, void *dst, const void *src, int64_t bytes);
int qlg_memcpy_d2h(qlg_ctx ctx, void *dst, const void *src, int64_t bytes);
int qlg_memcpy_d2d(qlg_ctx ctx, void *dst, const void *src, int64_t bytes);

/* ---- effective-Hamiltonian plans ----------------------------------------
 * Replaces the {L, W_l, W_r, R} operand vector and the function pointer
 * eff_ham_mul_two_site_state of
 *   include/qlmps/algorithm/vmps/lanczos_vmps_solver_impl.h:114-120, 260-289.
 * Index orders follow the reference: L[bra, mpo, ket], W[wl, s_in, s_out, wr],
 * R[bra, mpo, ket], psi[a, s1, s2, b].  The plan pairs quantum-number sectors
 * once and keeps the GEMM task lists for every later matvec with the same
 * operators; the operator data must stay alive and unchanged meanwhile.
 * The Krylov vector layout (block list and offsets) is the closure of psi's
 * blocks under H_eff; read it with qlg_plan_vector_layout.               */
int qlg_plan_two_site(qlg_ctx ctx, const qlg_bsten *L, const qlg_bsten *Wl,
                      const qlg_bsten *Wr, const qlg_bsten *R,
                      const qlg_bsten *psi, int precision, qlg_plan *out);
/* One-site form, lanczos_vmps_solver_impl.h:291-304 (psi[a, s, b]). */
int qlg_plan_one_site(qlg_ctx ctx, const qlg_bsten *L, const qlg_bsten *W,
                      const qlg_bsten *R, const qlg_bsten *psi,
                      int precision, qlg_plan *out);
/* DMRG-form (super-block) operator: replaces the SuperBlockHamiltonianTerms
 * argument of LanczosSolver(SuperBlockHamiltonianTerms<TenT>&, TenT*, const
 * LanczosParams&), include/qlmps/algorithm/dmrg/lanczos_dmrg_solver_impl.h:61-66,
 * whose matvec (:199-283) is
 *   H psi = sum_j [sum_i L_i (x) W_l(i, j)] psi [sum_k W_r(j, k) (x) R_k],
 * with the term groups of DMRGExecutor::SetEffectiveHamiltonianTerms_
 * (include/qlmps/algorithm/dmrg/dmrg_impl.h:319-341).  Group j holds
 * nleft[j] pairs {L_i, W_l(i, j)} and nright[j] pairs {W_r(j, k), R_k}; the
 * groups' pairs are stored back to back in left[] and right[].  Block
 * operators are rank 2, L_i[a (OUT), a' (IN)] and R_k[b (IN), b' (OUT)]; site
 * operators are W[s_in (IN), s_out (OUT)].  A block operator used by several
 * terms is passed as the same descriptor pointer (the reference shares it by
 * pointer).  The plan stacks the operators into MPO form once (bond elements
 * grouped by charge) and is then used like a two-site plan.               */
typedef struct {
  const qlg_bsten *op0, *op1;  /* {L_i, W_l(i,j)} or {W_r(j,k), R_k} */
} qlg_term;
int qlg_plan_super_block(qlg_ctx ctx, int64_t ngroups, const int64_t *nleft,
                         const qlg_term *left, const int64_t *nright,
                         const qlg_term *right, const qlg_bsten *psi,
                         int precision, qlg_plan *out);
int qlg_plan_destroy(qlg_plan plan);
int qlg_plan_set_precision(qlg_plan plan, int precision);
/* Algorithmic FLOPs of one matvec: sum of 2*M*N*K (8*M*N*K complex) over the
 * sector-GEMM pairs of the reference's four-step schedule
 * (lanczos_vmps_solver_impl.h:268-272).  exec_flops: FLOPs the plan's own
 * kernels issue (per product term).  Either pointer may be NULL.          */
int qlg_plan_flops(qlg_plan plan, double *algo_flops, double *exec_flops);
/* Vector layout: *nblocks, *size (elements); coords [nblocks*rank] and
 * offsets [nblocks] are filled when non-NULL (blocks in lexicographic order
 * of their sector coordinates, offsets = prefix sums of block sizes).     */
int qlg_plan_vector_layout(qlg_plan plan, int64_t *nblocks, int64_t *size,
                           int32_t *coords, int64_t *offsets);
/* Workspace bytes held by the plan (intermediates + task lists). */
int qlg_plan_workspace(qlg_plan plan, int64_t *bytes);
/* Kernel time of the last matvec per stage (ms, HIP events on the plan's
 * stream): [0] = first contraction, [1] = MPO stage, [2] = last contraction.
 * Valid after qlg_plan_enable_timing(plan, 1) and a synchronised matvec.  */
int qlg_plan_enable_timing(qlg_plan plan, int on);
int qlg_plan_stage_ms(qlg_plan plan, double *ms3);
/* FLOPs each stage issues per matvec (per product term), same order. */
int qlg_plan_stage_flops(qlg_plan plan, double *flops3);
/* Algorithmic work of each stage, same order: its share of the reference
 * schedule's FLOPs (the first contraction, the two MPO contractions, the
 * last contraction of lanczos_vmps_solver_impl.h:268-272) and its HBM bytes
 * with every operand read once and every result written once.  Either
 * pointer may be NULL.  No counterpart in the reference (roofline input). */
int qlg_plan_stage_work(qlg_plan plan, double *algo_flops3, double *algo_bytes3);

/* out = H_eff * psi, both in the plan's vector layout (device). */
int qlg_matvec(qlg_plan plan, const void *psi, void *out);

/* ---- contractions around the solve (same grouped-GEMM machinery) ------------
 * qlg_plan_contract: C = A x B over A's last k axes and B's first k axes
 * (result axes: A's free axes then B's), with optional conjugation of either
 * operand -- the reference's Contract(A, B, {{..}, {..}}, C) for the
 * products it forms around the solve (psi0 = A_l . A_r,
 * two_site_update_finite_vmps_impl.h:278-279; U.S and S.V^T, :322-331).
 * qlg_plan_env: the environment updates UpdateSiteLenvs (side 'l') and
 * UpdateSiteRenvs (side 'r'), include/qlmps/algorithm/vmps/vmps_init.h:343-373,
 * from (env, mps tensor, mpo tensor).  The result layout (blocks in
 * lexicographic order, prefix-sum offsets) is read with
 * qlg_plan_vector_layout; result axes are [Inverse(A.b), W.w1, A.b] for 'l'
 * and [Inverse(B.a), W.w0, B.a] for 'r'.
 * qlg_plan_transpose: the tensor with its axes in a new order, result axis
 * i = source axis perm[i] -- qlten's Transpose(axes) as the reference calls
 * it (two_site_update_finite_tdvp_impl.h:36); the DMRG form uses it to lay
 * block-operator groups out one operator after another.
 * qlg_execute runs such a plan: in0/in1 = (A, B), (env, mps) or (t, unused);
 * NULL reuses the data pointers given at planning time; out is a caller
 * buffer of the layout's size.                                             */
int qlg_plan_contract(qlg_ctx ctx, const qlg_bsten *A, const qlg_bsten *B, int k, int conj_a, int conj_b,
                      qlg_plan *out);
int qlg_plan_env(qlg_ctx ctx, const qlg_bsten *env, const qlg_bsten *mps, const qlg_bsten *mpo, char side,
                 qlg_plan *out);
int qlg_plan_transpose(qlg_ctx ctx, const qlg_bsten *t, const int32_t *perm, qlg_plan *out);
int qlg_execute(qlg_plan plan, const void *in0, const void *in1, void *out);

/* ---- Krylov vector kernels (flat vectors of n elements) ------------------
 * qlten LinearCombine / Normalize / Dag+Contract equivalents,
 * lanczos_vmps_solver_impl.h:141,155-161,173-183,238.  Reductions are
 * deterministic (fixed grid, ordered two-pass).  Scalars are host values;
 * complex scalars are (re, im) pairs.                                      */
int qlg_dot(qlg_ctx ctx, int dtype, int64_t n, const void *x, const void *y,
            double *out /* sum x*conj(y); 1 or 2 doubles */);
int qlg_nrm2(qlg_ctx ctx, int dtype, int64_t n, const void *x, double *out);
int qlg_scal(qlg_ctx ctx, int dtype, int64_t n, const double *alpha, void *x);
int qlg_axpy(qlg_ctx ctx, int dtype, int64_t n, const double *alpha,
             const void *x, void *y);
/* y <- beta*y + sum_i coefs[i] * bases[i]; bases is a host array of k device
 * pointers. */
int qlg_linear_combine(qlg_ctx ctx, int dtype, int64_t n, int64_t k,
                       const double *coefs, const void *const *bases,
                       const double *beta, void *y);
/* qlg_linear_combine that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory during the
 * call (code between hipGraph captures, tight serving loops, harnesses that
 * forbid hipMalloc).  qlg_linear_combine stages its coefficient and pointer
 * tables in device buffers it allocates per call; this entry point stages
 * them in a fixed array in the library's device image instead (24 KiB per
 * GPU, held until the process ends; its address is looked up when a context
 * is created on the device).  Limit: k <= 1024 basis vectors; a larger k is
 * refused with QLG_ERR_INVALID.  Same arguments, results and rounding as
 * qlg_linear_combine; returns after the work is complete.  All calls on one
 * GPU share the array: calls that overlap in time on one GPU (from several
 * threads or contexts) race and must be serialised by the caller.        */
int qlg_linear_combine_noalloc(qlg_ctx ctx, int dtype, int64_t n, int64_t k,
                               const double *coefs, const void *const *bases,
                               const double *beta, void *y);

/* ---- Krylov solvers ---------------------------------------------------------
 * qlg_lanczos_gs: LanczosSolver, lanczos_vmps_solver_impl.h:114-248 (plain
 * three-term recurrence, no reorthogonalisation, stop when
 * E_old - E_new < error, or the Krylov dimension reaches the vector size, or
 * m == max_iterations - 1).  psi_inout holds the initial state on entry and
 * the normalised ground state on return.
 * qlg_lanczos_expmv: LanczosExpmvSolver,
 * include/qlmps/algorithm/tdvp/lanczos_expmv_solver_impl.h:70-184
 * (psi <- exp(-i*tau*H_eff) psi, norm of psi kept; QLG_C128 plans only, in
 * the plan's precision).  */
int qlg_lanczos_gs(qlg_plan plan, void *psi_inout,
                   const qlg_lanczos_params *params, qlg_lanczos_result *res);
int qlg_lanczos_expmv(qlg_plan plan, void *psi_inout, double tau,
                      const qlg_lanczos_params *params, int64_t *iters);
/* How the two solvers above run.  Each iteration ends on the GPU: an FP64
 * kernel solves the tridiagonal problem (the lowest eigenvalue of T for
 * qlg_lanczos_gs; exp(-i*tau*T) e_1 for qlg_lanczos_expmv) and applies the
 * stop rule, and the host checks the device's stop flag once per chunk of
 * iterations (kernels queued past the stop return at once).  For
 * qlg_lanczos_gs the final Ritz vector comes from the same tridiagonal solve
 * on the host that TridiagGsSolver does; qlg_lanczos_expmv combines the
 * basis with the device's coefficients.  Krylov dimensions above 2047 (ground
 * state) or 64 (expmv) run the iteration-by-iteration host form instead.
 * qlg_plan_use_graphs(plan, 1) also applies to these solvers: each chunk of
 * iterations is captured as a hipGraph on first use and replayed by later
 * solves on the same plan with the same parameters.
 * qlg_plan_set_reorth(plan, passes): full reorthogonalisation of every new
 * residual against the whole Krylov basis by classical Gram-Schmidt, applied
 * `passes` times per iteration (0: off, the reference's plain three-term
 * recurrence, lanczos_vmps_solver_impl.h:173-183; 1 or 2).  The residual norm
 * b_m becomes the norm after reorthogonalisation.  No counterpart in the
 * reference; meant for the reduced-precision modes.                        */
int qlg_plan_set_reorth(qlg_plan plan, int passes);

/* ---- independent chains in one launch ------------------------------------------
 * The batching form of the path: several effective Hamiltonians (each a
 * two-site, one-site or super-block plan of its own chain and site) are solved
 * together.  qlg_plan_batch copies the member plans' operators into one plan
 * whose every matvec stage is a single grouped launch over all members (the
 * members may be destroyed afterwards).  Its vectors are the members' vectors
 * concatenated: member m occupies [offsets[m], offsets[m+1]) of a batch
 * vector, in the member's own layout (qlg_plan_batch_offsets).  qlg_matvec
 * applies all members at once.  qlg_lanczos_gs_batch runs LanczosSolver
 * (lanczos_vmps_solver_impl.h:114-248) for every member in lock step, each
 * with its own scalars, stop rule and result (res[m]); the iteration counts
 * and energies are those of separate solves.  qlg_plan_use_graphs(plan, 1)
 * captures each Lanczos iteration's device work as a hipGraph on first use
 * and replays it in later solves on the same plan.  Real (QLG_F64) FP64
 * members only.  No counterpart in the reference (which solves one site of
 * one chain at a time).                                                     */
int qlg_plan_batch(qlg_ctx ctx, int64_t nplans, const qlg_plan *plans, qlg_plan *out);
int qlg_plan_batch_offsets(qlg_plan plan, int64_t *nplans, int64_t *offsets /* nplans + 1, may be NULL */);
int qlg_plan_use_graphs(qlg_plan plan, int on);
int qlg_lanczos_gs_batch(qlg_plan plan, void *psi_inout, const qlg_lanczos_params *params,
                         qlg_lanczos_result *res);

/* ---- block SVD with truncation -------------------------------------------------
 * The qlten SVD of the reference's two-site updates,
 *   SVD(t, ldims, div, trunc_err, Dmin, Dmax, u, s, vt, &actual_trunc_err, &D)
 * (two_site_update_finite_vmps_impl.h:302-307, dmrg_impl.h:249-254,
 * two_site_update_finite_tdvp_impl.h:339-344).  t's first ldims axes are the
 * rows, the others the columns; ldiv (qn_len values) is the divergence the
 * left factor u keeps.  The new bond carries one sector per charge
 * q = ldiv - flow(left axes) with kept singular values, in ascending charge
 * order.  Truncation (this library's rule; qlten's is not part of the
 * reference): pool the singular values of all sectors and keep the smallest D
 * whose discarded weight sum(s^2)/total <= trunc_err, clamped to
 * [Dmin, Dmax]; actual_trunc_err is that discarded weight.  No value is
 * rounded to zero: kept singular values below ~1e-6 of their sector's largest
 * are resolved again on their own scale; vectors of exactly zero values are
 * completed to an orthonormal set.  Results live in the handle:
 *   u  [left axes..., bond (OUT)], same dtype as t;
 *   s  [bond (IN), bond (OUT)], diagonal blocks, always QLG_F64;
 *   vt [bond (IN), right axes...], same dtype as t;
 * each in the canonical layout (blocks in lexicographic order of their sector
 * coordinates, offsets = prefix sums).  qlg_svd_result gives D, the
 * truncation error and the bond's sector count; qlg_svd_bond its sectors;
 * qlg_svd_tensor (which: 0 u, 1 s, 2 vt) a tensor's block list and device
 * data, valid until qlg_svd_destroy.  Array arguments may be NULL to query
 * counts first.                                                              */
typedef struct qlg_svd_s *qlg_svd;
int qlg_block_svd(qlg_ctx ctx, const qlg_bsten *t, int32_t ldims, const int32_t *ldiv, double trunc_err,
                  int64_t Dmin, int64_t Dmax, qlg_svd *out);
int qlg_svd_result(qlg_svd r, double *trunc_err, int64_t *D, int64_t *bond_nsect);
int qlg_svd_bond(qlg_svd r, int32_t *qn /* nsect * qn_len */, int64_t *deg /* nsect */);
int qlg_svd_tensor(qlg_svd r, int which, int64_t *nblocks, int64_t *size, int32_t *coords, int64_t *offsets,
                   const void **data);
int qlg_svd_destroy(qlg_svd r);

/* ---- batched Hermitian eigen-decomposition ---------------------------------------
 * The solver qlg_block_svd runs on the Gram matrices of its sectors, exported
 * for callers with their own batches.  Group k holds batch matrices of order n, row-major and
 * Hermitian, at a (dtype QLG_F64 or QLG_C128); on return row j of each matrix
 * holds eigenvector j (complex-conjugated for QLG_C128) and w[0..n) its
 * eigenvalues in ascending order.  Divide and conquer; groups run
 * concurrently.  The call returns after the work is complete (it checks the
 * solver's convergence flags); QLG_ERR_NUMERIC if a matrix did not converge. */
typedef struct {
  int64_t n;      /* order of each matrix   */
  int64_t batch;  /* matrices in the group  */
  void *a;        /* [batch][n][n] device   */
  double *w;      /* [batch][n] device      */
} qlg_eigh_group;
int qlg_eigh_batched(qlg_ctx ctx, int dtype, int64_t ngroups, const qlg_eigh_group *groups);

/* ---- host helpers (FP64) ------------------------------------------------------
 * qlg_tridiag_gs: TridiagGsSolver, lanczos_vmps_solver_impl.h:306-355
 * (lowest eigenpair of the symmetric tridiagonal T(a[0..n), b[0..n-1));
 * vec may be NULL for jobz 'N').
 * qlg_tridiag_expme1: TridiagExpme1Solver, lanczos_expmv_solver_impl.h:199-240
 * (exp(-i*tau*T) e_1, res = n interleaved complex values).               */
int qlg_tridiag_gs(const double *a, const double *b, int64_t n, char jobz,
                   double *eng, double *vec);
int qlg_tridiag_expme1(const double *a, const double *b, int64_t n,
                       double tau, double *res);

#ifdef __cplusplus
}
#endif
#endif /* QLGPU_H */
