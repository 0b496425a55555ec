/*
 * itx.h -- C ABI of the MI355X (gfx950) DMRG H*psi engine.
 *
 * Every entry point takes plain pointers and sizes (no C++ or torch types),
 * returns an int status (ITX_OK == 0) and is stream-ordered on the context's
 * HIP stream.  Device pointers are never freed by the library unless they
 * came from itx_malloc.  A context is thread-compatible, not thread-safe.
 *
 * The functions replace the numerical layer underneath ITensor's storage
 * plug-ins (the BLAS/LAPACK call sites of the reference):
 *   - dense contraction      : contract() in tensor/contract.h:26-33,
 *                              tensor/contract.cc:843-866 (permute + dgemm
 *                              via gemm_wrapper, tensor/lapack_wrap.cc:116-161)
 *   - block-sparse contraction: doTask(Contract&,QDense,QDense,ManageStore&)
 *                              in itdata/qdense.cc:671-743, which runs one
 *                              contract() per block pair from
 *                              getContractedOffsets (itdata/qutil.h:93-242)
 *   - vector kernels         : daxpy/dnrm2/ddot/dscal wrappers,
 *                              tensor/lapack_wrap.cc:26-75,353
 *   - symmetric eigensolver  : dsyev_wrapper, tensor/lapack_wrap.cc:323,
 *                              as called by diagHermitian
 *                              (tensor/algs_impl.h:91-139)
 *   - permuted add           : the transform(permute(B),A,Adder) of
 *                              PlusEQ (itdata/dense.cc:362-415,
 *                              itdata/qdense.cc:506-550)
 *
 * Layout convention: tensors are column-major (first index fastest), as
 * ITensor's Range (tensor/range.h:196-206).  Contraction labels follow
 * computeLabels (tensor/contract.h:155-202): a negative label marks a
 * contracted index (same label on A and B), positive labels name the
 * uncontracted indices, and C's labels give the order of C's indices.
 */
#ifndef ITX_H
#define ITX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define ITX_OK 0
#define ITX_ERR_ARG 1         /* invalid argument (shape/label mismatch) */
#define ITX_ERR_HIP 2         /* HIP runtime error                       */
#define ITX_ERR_NOMEM 3       /* device allocation failed                */
#define ITX_ERR_UNSUPPORTED 4 /* size or form outside what is built      */
#define ITX_ERR_NOCONV 5      /* eigensolver did not converge            */

/* ---- arithmetic forms of the contraction engine ----------------------- */
#define ITX_PREC_F64 0    /* FP64 MFMA, FP64 accumulate (the plain form)  */
#define ITX_PREC_F32 1    /* FP32 MFMA, FP32 accumulate                   */
#define ITX_PREC_BF16X3 2 /* 3-product BF16 split, FP32 accumulate        */

#define ITX_MAX_RANK 12

typedef struct itx_ctx itx_ctx;
typedef struct itx_plan itx_plan;
typedef struct itx_chain itx_chain;

/* ---- context, stream, memory ---------------------------------------- */
/* Create a context on HIP device `device` using the null stream. */
int itx_create(int device, itx_ctx **out);
int itx_destroy(itx_ctx *ctx);
/* Make every later call on ctx run on `stream` (a hipStream_t). */
int itx_set_stream(itx_ctx *ctx, void *stream);
void *itx_get_stream(itx_ctx *ctx);
int itx_synchronize(itx_ctx *ctx);
/* Text of the last error on ctx (or of the last failed itx_create). */
const char *itx_last_error(itx_ctx *ctx);
/* Number of compute units of the context's device (0 on error). */
int itx_device_cus(itx_ctx *ctx);

/* Stream-ordered caching device allocator (blocks are reused, never
 * returned to HIP before itx_destroy or itx_trim). */
int itx_malloc(itx_ctx *ctx, size_t bytes, void **ptr);
int itx_free(itx_ctx *ctx, void *ptr);
int itx_trim(itx_ctx *ctx);
int itx_memcpy_htod(itx_ctx *ctx, void *dst, const void *src, size_t bytes);
int itx_memcpy_dtoh(itx_ctx *ctx, void *dst, const void *src, size_t bytes);
int itx_memcpy_dtod(itx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Pinned host memory from the context's caching pool, and a copy in any
 * direction ordered on the context's stream that does not wait (between
 * pinned host and device memory, or device and device).  Blocks return to
 * the pool in stream order: a block freed and handed out again is only
 * touched by copies enqueued after the ones that used it.  The host must
 * not read a block a copy writes before the stream has passed it.  Used to
 * keep inactive DMRG environments in host memory, where the reference
 * writes them to disk (LocalMPO::setLHlim / setRHlim, mps/localmpo.h:
 * 620-681). */
int itx_host_alloc(itx_ctx *ctx, size_t bytes, void **ptr);
int itx_host_free(itx_ctx *ctx, void *ptr);
int itx_memcpy_async(itx_ctx *ctx, void *dst, const void *src, size_t bytes);
/* x[i] = v for i < n (a fill kernel, safe under graph capture). */
int itx_fill(itx_ctx *ctx, double *x, size_t n, double v);

/* ---- dense contraction ------------------------------------------------
 * C = alpha * contract(A, B) + beta * C, with the label convention above.
 * Replaces contract(TenRefc A, Labels ai, TenRefc B, Labels bi, TenRef C,
 * Labels ci, alpha, beta) (tensor/contract.h:26-33, contract.cc:843-866)
 * and its CProps permute->gemm->permute plan (contract.cc:137-824): the
 * permutations are folded into the tile loads and the tile stores of one
 * MFMA GEMM kernel.  prec selects the arithmetic form (ITX_PREC_*). */
int itx_contract_dense(itx_ctx *ctx,
                       const double *A, int rA, const int64_t *dimsA, const int32_t *labA,
                       const double *B, int rB, const int64_t *dimsB, const int32_t *labB,
                       double *C, int rC, const int64_t *dimsC, const int32_t *labC,
                       double alpha, double beta, int prec);

/* ---- block-sparse (QN) contraction ------------------------------------
 * Executes the block-pair list of getContractedOffsets
 * (itdata/qutil.h:93-242) as one persistent grouped-GEMM launch, with the
 * first-touch beta semantics of doTask(Contract,QDense,QDense)
 * (itdata/qdense.cc:700-731): every C block named by a task is
 * overwritten with alpha * (sum over its tasks of A_blk * B_blk); C blocks
 * named by no task are left untouched.
 * Task t contracts the A block at element offset offs[3t] (dims
 * blkdimsA[t*rA .. t*rA+rA-1]) with the B block at offs[3t+1] (dims
 * blkdimsB[t*rB ..]) into the C block at offs[3t+2], whose dims follow from
 * the labels.  Labels are shared by all tasks. */
int itx_contract_qn(itx_ctx *ctx,
                    const double *A, int rA, const int32_t *labA,
                    const double *B, int rB, const int32_t *labB,
                    double *C, int rC, const int32_t *labC,
                    int ntasks, const int64_t *offs,
                    const int64_t *blkdimsA, const int64_t *blkdimsB,
                    double alpha, int prec);

/* As itx_contract_qn, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid it).  The call's task list, offset tables
 * and split-K partial tiles go into a static device array of the library,
 * 32 MiB per device, which each device holds until the process ends; a
 * contraction whose plan needs more returns ITX_ERR_UNSUPPORTED (the
 * message names both sizes).  The tile comes from the tuning table or the
 * cost model, never from the autotuner.  Every call on a device shares
 * that device's array: calls on one stream are ordered, calls that overlap
 * on one device (other contexts or streams) race.  Nothing in the array
 * needs to start at zero. */
int itx_contract_qn_noalloc(itx_ctx *ctx,
                            const double *A, int rA, const int32_t *labA,
                            const double *B, int rB, const int32_t *labB,
                            double *C, int rC, const int32_t *labC,
                            int ntasks, const int64_t *offs,
                            const int64_t *blkdimsA, const int64_t *blkdimsB,
                            double alpha, int prec);

/* ---- reusable plans (no host work or uploads at execution) ------------
 * A plan holds the device task list of one contraction structure; it is
 * executed any number of times on new data (e.g. the three H*psi products
 * of one Davidson call) and may be captured in a hipGraph. */
int itx_plan_dense(itx_ctx *ctx,
                   int rA, const int64_t *dimsA, const int32_t *labA,
                   int rB, const int64_t *dimsB, const int32_t *labB,
                   int rC, const int64_t *dimsC, const int32_t *labC,
                   int prec, itx_plan **out);
int itx_plan_qn(itx_ctx *ctx,
                int rA, const int32_t *labA, int rB, const int32_t *labB,
                int rC, const int32_t *labC,
                int ntasks, const int64_t *offs,
                const int64_t *blkdimsA, const int64_t *blkdimsB,
                int prec, itx_plan **out);
int itx_plan_execute(itx_ctx *ctx, const itx_plan *plan,
                     const double *A, const double *B, double *C,
                     double alpha, double beta);
/* Tile autotuner.  With autotuning on, a plan whose size class (precision,
 * operand layout, log2 buckets of block shape, summed K and block count)
 * has no entry in the context's tuning table times every MFMA tile shape
 * on scratch buffers and records the fastest; later plans of that class
 * use the table.  Plans must not be created under graph capture while
 * autotuning is on.  The table can be saved and loaded as text. */
int itx_set_autotune(itx_ctx *ctx, int on);
int itx_tune_save(itx_ctx *ctx, const char *path);
int itx_tune_load(itx_ctx *ctx, const char *path);
int itx_tune_size(itx_ctx *ctx);

/* Pre-split operands of the BF16x3 form (no counterpart in the reference):
 * itx_split_bf16 writes y[i] = hi | lo << 16, the two BF16 terms of x[i]
 * rounded to FP32 (hi = bf16(x), lo = bf16(x - hi)) that the BF16x3
 * kernels otherwise form at every tile load; itx_plan_set_split(plan, 1)
 * makes a BF16x3 plan whose operands are both contiguous along the
 * contracted index read such words (A and B passed to itx_plan_execute
 * then point to the words, n per operand element), so an operand reused
 * by many products (the environments of a Davidson call) is split once and
 * streamed at 4 bytes per element.  The results equal the plan's without
 * pre-splitting bit for bit.  ITX_ERR_UNSUPPORTED for other plans. */
int itx_split_bf16(itx_ctx *ctx, const double *x, uint32_t *y, size_t n);
int itx_plan_set_split(itx_ctx *ctx, itx_plan *plan, int on);

/* Algorithmic FP64-equivalent flops of one execution (sum of 2*m*n*k). */
double itx_plan_flops(const itx_plan *plan);
/* What a plan runs (diagnostic; no counterpart in the reference): the
 * kernel's tile index (ITX_FORCE_TILE numbering; 5 = skinny, 10 = small
 * blocks), the arithmetic form its kernel computes in (ITX_PREC_*: the
 * HBM-bound skinny path computes in FP64 in every form), the number of
 * block-pair tasks and of work units.  Any output pointer may be NULL. */
int itx_plan_info(const itx_plan *plan, int *tile, int *form, int *ntasks, int *nunits);
int itx_plan_destroy(itx_ctx *ctx, itx_plan *plan);

/* ---- H*psi chains replayed as a HIP graph -------------------------------
 * A chain is a fixed sequence of plan executions, stage i computing
 * C[i] = alpha[i] * contract(A[i], B[i]) + beta[i] * C[i] (the stages of
 * LocalOp::product, mps/localop.h:346-364, whose structure stays fixed
 * across the products of one Davidson call, iterativesolvers.h:160-330).
 * The first itx_chain_launch builds one hipGraph with a kernel node per
 * stage (each after the previous one) and instantiates it; later launches
 * patch the operand pointers and scalars of the nodes whose values changed
 * and replay the graph on the context's stream.  alpha / beta may be NULL
 * (1 and 0).  The chain does not own the plans: they must outlive it. */
int itx_chain_create(itx_ctx *ctx, int nstages, itx_plan *const *plans, itx_chain **out);
int itx_chain_launch(itx_ctx *ctx, itx_chain *chain, const double *const *A, const double *const *B,
                     double *const *C, const double *alpha, const double *beta);
int itx_chain_destroy(itx_ctx *ctx, itx_chain *chain);

/* ---- vector kernels (FP64, wavefront reductions) ----------------------
 * Results of reductions are written to host memory (the call waits for
 * the stream) by the plain forms, and to device memory by the _dev forms
 * (no wait).  dnrm2/ddot/daxpy/dscal of tensor/lapack_wrap.cc:26-75,353. */
int itx_dot(itx_ctx *ctx, const double *x, const double *y, size_t n, double *result);
int itx_nrm2(itx_ctx *ctx, const double *x, size_t n, double *result);
int itx_dot_dev(itx_ctx *ctx, const double *x, const double *y, size_t n, double *dresult);
/* out[j] = <V_j, y> for j < nv; V_j are nv device vectors of length n. */
int itx_multi_dot(itx_ctx *ctx, const double *const *V, int nv,
                  const double *y, size_t n, double *out);
int itx_axpy(itx_ctx *ctx, double a, const double *x, double *y, size_t n);
int itx_scal(itx_ctx *ctx, double a, double *x, size_t n);
/* y = sum_j c[j] * V_j (y may not alias any V_j). */
int itx_lincomb(itx_ctx *ctx, const double *const *V, const double *c, int nv,
                double *y, size_t n);

/* ---- permuted add ------------------------------------------------------
 * Y += alpha * permute(X): Y has dims dimsY (rank r); element Y[i_0..]
 * receives X at the position where X's index perm[j] equals Y's index j,
 * i.e. X dims satisfy dimsX[perm[j]] == dimsY[j].  Covers PlusEQ
 * (itdata/dense.cc:362-415) and the block loop of QDense PlusEQ
 * (itdata/qdense.cc:506-550), and a plain permuted copy when beta == 0. */
int itx_permute(itx_ctx *ctx, const double *X, double *Y, int r,
                const int64_t *dimsY, const int32_t *perm, double alpha, double beta);

/* ---- batched strided block copy ----------------------------------------
 * For each of ntasks tasks t, with desc_t = desc + t*(2+3r) holding
 * [src_off, dst_off, dims[r], src_strides[r], dst_strides[r]]:
 *   Y[dst_off + sum_j i_j*dst_strides[j]] =
 *       alpha * X[src_off + sum_j i_j*src_strides[j]] + beta * Y[...]
 * over all multi-indices i < dims (beta == 0 never reads Y).  One launch
 * moves every block of a QDense PlusEQ (itdata/qdense.cc:506-550, incl. the
 * block-set widening of :582-663) or of a QCombiner combine / uncombine
 * (itdata/qcombiner.cc:125-300).  Tasks must not write overlapping
 * elements. */
int itx_block_copy(itx_ctx *ctx, const double *X, double *Y, int ntasks, int r,
                   const int64_t *desc, double alpha, double beta);

/* ---- symmetric eigensolver (FP64) --------------------------------------
 * Eigen-decomposition of the symmetric n x n column-major matrix A (device):
 * on return w holds the eigenvalues in DESCENDING order and A the matching
 * orthonormal eigenvectors as columns.  This is what diagHermitian obtains
 * from dsyev on -M (tensor/algs_impl.h:122-135, tensor/algs.cc:34-41). */
int itx_syev(itx_ctx *ctx, int n, double *A, double *w);
/* nmat independent problems in one launch: matrix b (n[b] x n[b]) at
 * A + a_offs[b], eigenvalues to w + w_offs[b].  The per-block
 * diagonalisation of a block-diagonal QN density matrix
 * (hermitian.cc:206-257). */
int itx_syev_batched(itx_ctx *ctx, int nmat, const int32_t *n, const int64_t *a_offs, double *A,
                     const int64_t *w_offs, double *w);
/* As itx_syev_batched, and the same eigenvalues (bit for bit) also in host
 * memory, w_host + w_offs[b], when the call returns -- while the
 * eigenvectors are still being completed on the stream.  The truncation
 * that follows the per-block diagonalisation in diagHermitian
 * (hermitian.cc:259-307) needs only the eigenvalues, so the host can decide
 * the kept dimensions without waiting for the eigenvector kernels. */
int itx_syev_batched_host(itx_ctx *ctx, int nmat, const int32_t *n, const int64_t *a_offs, double *A,
                          const int64_t *w_offs, double *w, double *w_host);

/* ---- singular value decomposition (FP64) --------------------------------
 * A (m x n, column-major, device, not modified) = U diag(S) V^T with
 * k = min(m,n): U is m x k and V is n x k (column-major, orthonormal
 * columns), S holds the k singular values in DESCENDING order.  Columns
 * of the long side that belong to an exactly zero singular value are set
 * to zero.  Replaces SVD(M,U,D,V) as svdImpl calls it per matrix block
 * (svd.cc:87,206-213 -> SVDRef, tensor/algs.cc:586-631). */
int itx_gesvd(itx_ctx *ctx, int m, int n, const double *A, double *U, double *S, double *V);
/* nmat independent problems in one launch (the blocks of a QN tensor):
 * matrix b (m[b] x n[b]) at A + a_offs[b], factors at U + u_offs[b],
 * S + s_offs[b], V + v_offs[b]. */
int itx_gesvd_batched(itx_ctx *ctx, int nmat, const int32_t *m, const int32_t *n, const int64_t *a_offs,
                      const double *A, const int64_t *u_offs, double *U, const int64_t *s_offs, double *S,
                      const int64_t *v_offs, double *V);

#ifdef __cplusplus
}
#endif
#endif /* ITX_H */
