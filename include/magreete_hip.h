/* C ABI of libmagreete_hip.so: the coupled-dipole hot path on MI355X (gfx950).
 *
 * Conventions
 *  - All buffers are device pointers allocated by the caller (torch owns
 *    them); matrices are row-major complex (interleaved re, im) with a
 *    leading dimension in elements.  Points are row-major float64 (x, y[, z]).
 *  - Row d*m + a, column d*j + b: the reference's (M, N, d, d) Green tensor
 *    transposed to (d M, d N) (Transmission3D.py:267, Transmission2D.py:279).
 *  - Every call is asynchronous on `stream` (a hipStream_t; NULL = default
 *    stream) and returns an int status: 0 ok, < 0 invalid argument / runtime
 *    failure (text from mg_last_error()).  LU info (first exactly-zero pivot,
 *    1-based) is written to a device int, as LAPACK's info.
 *  - A handle owns the LU workspace; one handle must not be used from two
 *    streams at once.
 */
#ifndef MAGREETE_HIP_H
#define MAGREETE_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mg_handle mg_handle;

/* Green's-function kinds (one per reference class) */
enum { MG_2D_TM = 0, MG_2D_TE = 1, MG_3D_VECTOR = 2, MG_3D_SCALAR = 3 };
/* element precisions: complex128; complex64; complex64 with products from
 * two-part bf16 splits (three bf16 MFMA products, ~16 bits) or three-part
 * splits (six products, ~24 bits: FP32-level) */
enum { MG_C128 = 0, MG_C64 = 1, MG_BF16X3 = 2, MG_BF16X6 = 3 };
/* assembly flags */
enum {
  MG_F_REGULARIZE = 1, MG_F_ZERO_SELF = 2, MG_F_DIAG = 4, MG_F_NANFILL = 8,
  MG_F_PERIODIC_X = 16, MG_F_PERIODIC_Y = 32, MG_F_PERIODIC_Z = 64,
  MG_F_BESSEL_EXACT = 128 /* 2D: accurate J/Y instead of the reference's torch values */
};
/* status codes */
enum { MG_OK = 0, MG_ERR_ARG = -1, MG_ERR_HIP = -2, MG_ERR_KIND = -3, MG_ERR_NOMEM = -4, MG_ERR_SYNC = -5,
       MG_ERR_UNSUPPORTED = -6 };

int mg_version(void);
const char* mg_last_error(void);
int mg_create(mg_handle** out);
int mg_destroy(mg_handle* h);

/* Green's assembly: out[d i + a, d j + b] = scale * G_ab(rows[i] - cols[j]).
 * Replaces the Green's functions and G0 of the reference
 * (Transmission2D.py:63-91, 263-280, 595-618, 777-785;
 *  Transmission3D.py:47-77, 240-268, 573-601, 751-777) and the construction
 * of M = I - k0^2 alpha G with its self-interaction shift
 * (Transmission2D.py:222-228, 733-739; Transmission3D.py:197-203, 709-715)
 * through the flags: ZERO_SELF zeroes block (i, i); DIAG sets entry (r, r)
 * to diag; NANFILL replaces NaN entries by fill (LDOS fix-up,
 * Transmission3D.py:353-362); REGULARIZE and PERIODIC_* as the reference's
 * regularize / periodic arguments.  Either output may be NULL (c128 and/or a
 * complex64 copy written in the same pass). */
int mg_assemble(int kind, const double* rows, int64_t nrows, const double* cols, int64_t ncols, double k0,
                double scale_re, double scale_im, double diag_re, double diag_im, double fill_re, double fill_im,
                int flags, double radius, void* out_c128, int64_t ld_c128, void* out_c64, int64_t ld_c64,
                void* stream);

/* C = alpha A B + beta C, row-major complex (prec MG_C128, MG_C64, MG_BF16X3, MG_BF16X6). */
int mg_cgemm(int prec, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im, const void* A,
             int64_t lda, const void* B, int64_t ldb, double beta_re, double beta_im, void* C, int64_t ldc,
             void* stream);

/* mg_cgemm for tall-skinny products (the residual and Krylov products
 * M V of Transmission3D.py:205 / Transmission2D.py:230's systems when solved
 * iteratively): K is split over a workspace of mg_cgemm_work_elems(prec, m,
 * n, k) elements of C's type so that the tile grid fills the CUs, and the
 * partial products are summed in a fixed order.  work may be NULL (or the
 * count 0), which runs mg_cgemm; MG_BF16X3 / MG_BF16X6 never split. */
int64_t mg_cgemm_work_elems(int prec, int64_t m, int64_t n, int64_t k);
int mg_cgemm_ws(int prec, int64_t m, int64_t n, int64_t k, double alpha_re, double alpha_im, const void* A,
                int64_t lda, const void* B, int64_t ldb, double beta_re, double beta_im, void* C, int64_t ldc,
                void* work, int64_t work_elems, void* stream);

/* LU with partial pivoting in place; the factor step of torch.linalg.solve
 * (Transmission2D.py:230, 742; Transmission3D.py:205, 717).  ipiv: n device
 * ints (0-based row swapped with row k); info: one device int. */
int mg_getrf(mg_handle* h, int prec, void* A, int64_t n, int64_t lda, int* ipiv, int* info, void* stream);
/* Multi-RHS solve from getrf's factors, B (n x nrhs) overwritten by X: the
 * solve step of torch.linalg.solve at the same call sites. */
int mg_getrs(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, const int* ipiv, void* B, int64_t ldb,
             int64_t nrhs, void* stream);
/* Stored inverses of the 256 x 256 diagonal blocks of L and U, so that the
 * solve runs as GEMMs only (repeated solves against a resident factor; same
 * call sites as mg_getrs).  dinv holds mg_dinv_elems(n) elements of the
 * factor's type; work holds n x nrhs elements. */
int64_t mg_dinv_elems(int64_t n);
int mg_getrf_dinv(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, void* dinv, void* stream);
int mg_getrs_dinv(mg_handle* h, int prec, const void* LU, int64_t n, int64_t lda, const int* ipiv, const void* dinv,
                  void* B, int64_t ldb, int64_t nrhs, void* work, void* stream);
/* d x d diagonal blocks of M^{-1} for the sites listed in `sites` (int64,
 * device; block s covers rows and columns d*sites[s] .. d*sites[s]+d-1),
 * from getrf's complex128 factors and mg_getrf_dinv's inverses.  Replaces
 * the full inverse W = solve(M, eye) that batch_LDOS_measurements forms
 * (Transmission2D.py:416, 911; Transmission3D.py:403, 916) when the LDOS is
 * measured at the scatterers themselves: there every term of
 * diag(G W G^T) reduces to W's diagonal blocks.  LU is overwritten by
 * U^{-1} and L^{-1} (the factor is consumed); a site index outside
 * [0, n / d) gives a NaN block.  colperm (int32, device, n):
 * the position ipiv's row swaps move each row to.  out: nsites * d * d
 * complex128, row-major blocks; work: mg_inverse_diag_work_elems(n)
 * complex128.  prec must be MG_C128. */
int64_t mg_inverse_diag_work_elems(int64_t n);
int mg_inverse_diag_blocks(mg_handle* h, int prec, void* LU, int64_t n, int64_t lda, const void* dinv,
                           const int* colperm, int d, const int64_t* sites, int64_t nsites, void* out, void* work,
                           int64_t work_elems, void* stream);
/* Block LDU factorisation without pivoting, in place (complex float; prec
 * MG_C64, MG_BF16X3 or MG_BF16X6): A = Lh D Uh with 256 x 256 diagonal blocks; dinv
 * (mg_ldu_factor_elems(prec, n) elements, 16-byte aligned) receives the
 * inverses of the D blocks (the first mg_ldu_dinv_elems(n)) and, for the
 * split precisions, the off-diagonal panels split into bf16 parts; A receives
 * the scaled off-diagonal blocks; work holds mg_ldu_work_elems(n) elements;
 * info: first exactly-zero pivot or 0.  The preconditioner of the
 * mixed-precision refinement (no counterpart in the reference).
 * mg_getrs_ldu solves with it by GEMMs only (work:
 * mg_ldu_solve_work_elems(prec, n, nrhs) elements, 16-byte aligned); the
 * split precisions read dinv alone, so their LDU may be null (the factorised
 * matrix need not be kept). */
int64_t mg_ldu_dinv_elems(int64_t n);
int64_t mg_ldu_factor_elems(int prec, int64_t n);
int64_t mg_ldu_work_elems(int64_t n);
int64_t mg_ldu_solve_work_elems(int prec, int64_t n, int64_t nrhs);
int mg_getrf_ldu(mg_handle* h, int prec, void* A, int64_t n, int64_t lda, void* dinv, void* work, int* info,
                 void* stream);
int mg_getrs_ldu(mg_handle* h, int prec, const void* LDU, int64_t n, int64_t lda, const void* dinv, void* B,
                 int64_t ldb, int64_t nrhs, void* work, void* stream);
/* Cap on the workgroups of one LU leaf launch of this handle (g <= 0: the CU
 * count).  The leaf synchronises its workgroups through a grid barrier, so
 * factorisations that run at the same time on one device (one handle each)
 * should keep the sum of their caps within the CU count. */
int mg_set_leaf_blocks(mg_handle* h, int g);
/* Waits for `stream`; MG_ERR_SYNC if a grid barrier of an LU leaf failed. */
int mg_check(mg_handle* h, void* stream);
/* Measurement hooks with no reference counterpart: HIP-event timing of the
 * LU trailing-update launches (pivoted LU: k_zgemm_dma_trail / k_cgemm_trail;
 * block LDU: k_ldu_update_dma / k_ldu_update) on the launching stream.
 * read returns launches, algorithmic flops (8 m n k each) and summed ms since
 * the last read, and synchronises the stream; read_classes returns the same
 * per kernel class in arrays of two (element 0: operands staged through
 * registers, element 1: operands DMA'd to LDS). */
int mg_profile_enable(mg_handle* h, int on);
int mg_profile_read(mg_handle* h, void* stream, long long* launches, double* flops, double* ms);
int mg_profile_read_classes(mg_handle* h, void* stream, long long* launches, double* flops, double* ms);
/* Element conversion of a rows x cols block; dir 0: c128 -> c64, 1: c64 -> c128. */
int mg_convert(int dir, const void* src, int64_t lds, void* dst, int64_t ldd, int64_t rows, int64_t cols,
               void* stream);

/* Fused propagation E = alpha G(meas, scat) Ek + beta E with G evaluated in
 * kernel (the matmul of propagate / propagate_ss: Transmission2D.py:190,
 * 245, 701, 760; Transmission3D.py:163, 220, 675, 733).  Only REGULARIZE and
 * PERIODIC_* flags apply. */
int mg_propagate(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0,
                 double alpha_re, double alpha_im, int flags, double radius, const void* Ek, int64_t ld_ek,
                 int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e, void* stream);
/* mg_propagate with the scatterers split over a workspace of
 * mg_propagate_work_bytes(kind, n_meas, n_scat, nrhs) bytes (16-byte
 * aligned; 0: no split), so that a grid of few output tiles still gives every
 * CU two workgroups; the parts are added in a fixed order. */
int64_t mg_propagate_work_bytes(int kind, int64_t n_meas, int64_t n_scat, int64_t nrhs);
int mg_propagate_ws(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0,
                    double alpha_re, double alpha_im, int flags, double radius, const void* Ek, int64_t ld_ek,
                    int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e, void* work,
                    int64_t work_bytes, void* stream);
/* mg_propagate_ws with the incident field evaluated in the kernels' final
 * pass instead of read from E: E = alpha G(meas, scat) Ek + E0 where E0 is
 * the plane wave (src_type 0) or Gaussian beam of waist w (1) of every
 * column, as mg_sources builds it (dirs: T x 2 of (cos(-theta),
 * sin(-theta)) in 2D, T x 3 directions in 3D; pol: T x d polarisations for
 * the vector kinds, else ignored).  E is write-only.  Replaces
 * generate_source at the measurement points followed by propagate
 * (Transmission3D.py:79-145, 147-179; Transmission2D.py:93-174, 176-205)
 * when the caller needs no E0 of its own. */
int mg_propagate_src(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat, double k0,
                     double alpha_re, double alpha_im, int flags, double radius, const void* Ek, int64_t ld_ek,
                     int64_t nrhs, int src_type, const double* dirs, const double* pol, double w, void* E,
                     int64_t ld_e, void* work, int64_t work_bytes, void* stream);
/* mg_propagate_ws without allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that forbid hipMalloc).  The split's partial sums go to a fixed device
 * array of mg_propagate_noalloc_max_bytes() bytes (64 MiB) that the library
 * holds on every device it runs on until the process ends; a call whose
 * mg_propagate_work_bytes exceeds it returns MG_ERR_ARG.  Same kernels,
 * split and summation order as mg_propagate_ws, hence the same result.
 * Calls that overlap on one device share the array and race. */
int64_t mg_propagate_noalloc_max_bytes(void);
int mg_propagate_noalloc(int kind, const double* meas, int64_t n_meas, const double* scat, int64_t n_scat,
                         double k0, double alpha_re, double alpha_im, int flags, double radius, const void* Ek,
                         int64_t ld_ek, int64_t nrhs, double beta_re, double beta_im, void* E, int64_t ld_e,
                         void* stream);

/* H_nu^(1)(x) for nu in {0, 1} (torch_hankel1, Transmission2D.py:46-61, 578-593);
 * exact = 0: the reference's torch values, 1: accurate to ~1e-15. */
int mg_hankel1(int nu, int exact, const double* x, int64_t n, void* out, void* stream);

/* out[c] = sum_r Y[r, c] X[r, c] without conjugation: the diagonal of
 * G_p W G_p^T once X = M^{-1} G_p^T (einsum 'ij,ji->i' at
 * Transmission3D.py:365 and siblings). */
int mg_coldot(const void* Y, int64_t ldy, const void* X, int64_t ldx, int64_t n, int64_t ncols, void* out,
              void* stream);

/* Plane waves (type 0) and Gaussian beams (type 1) at points
 * (generate_source, Transmission2D.py:93-143, 620-652;
 *  Transmission3D.py:86-120, 610-633).  space 2: dirs[t] = (cos(-theta),
 * sin(-theta)); space 3: dirs[t] = u_t.  pol NULL (d = 1): out is
 * npts x T; else out[(m d + a) T + t] = f_t(x_m) pol[t d + a]. */
int mg_sources(int space, int type, const double* points, int64_t npts, const double* dirs, int64_t T,
               const double* pol, int64_t d, double k0, double w, void* out, void* stream);

/* Gram-Schmidt passes of the mixed-precision refinement (no counterpart in
 * the reference, whose solve is a single torch.linalg.solve).  V_i = V + i sV
 * are n x T complex128 matrices (row stride ldv), one Krylov process per
 * column t.  mg_mdot: out[i ldo + t] = sum_r conj(V_i[r, t]) w[r, t] for
 * i < k, summed in a fixed order (work: mg_mdot_work_elems(k, T) complex128).
 * mg_maxpy: w[r, t] += alpha sum_i V_i[r, t] h[i ldh + t]. */
int64_t mg_mdot_work_elems(int64_t k, int64_t T);
int mg_mdot(const void* V, int64_t sV, int64_t ldv, int k, const void* w, int64_t ldw, int64_t n, int64_t T,
            void* work, void* out, int64_t ldo, void* stream);
/* Arnoldi column j of the GMRES least-squares problems of all T columns (no
 * reference counterpart): H ((m+1) x m x T complex128, entry (i, j) of column
 * t at H[(i m + j) T + t]) holds the Gram-Schmidt coefficients of column j;
 * hn (T doubles) the norms of the new vectors, which become h_{j+1,j}.  The
 * unrotated column is copied to H0, the j earlier Givens rotations (cs, sn:
 * m x T) and a new one are applied, g ((m+1) x T) is rotated, and est[0]
 * receives max_t |g_{j+1,t}| / bnorm[t] (device pointers, stream-ordered). */
int mg_givens(void* H, void* H0, void* cs, void* sn, void* g, const double* hn, const double* bnorm, int m, int j,
              int64_t T, double* est, void* stream);
/* The GMRES coefficients of all T columns after k Arnoldi steps (no
 * reference counterpart): y (k x T complex128, y[i T + t]) solves the upper-
 * triangular R_t y_t = g_t with R_t the first k x k of the rotated H of
 * mg_givens (layout as there) and g_t its first k entries. */
int mg_hess_solve(const void* H, const void* g, int m, int k, int64_t T, void* y, void* stream);
int mg_maxpy(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, double alpha, void* w,
             int64_t ldw, int64_t n, int64_t T, void* stream);
/* mg_maxpy followed by the column norms of the updated w (no reference
 * counterpart): norms[t] = ||w[:, t]||_2 (T doubles, summed in a fixed
 * order); work: mg_mdot_work_elems(1, T) complex128 elements. */
int mg_maxpy_norm(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, double alpha, void* w,
                  int64_t ldw, int64_t n, int64_t T, void* work, double* norms, void* stream);
/* w[r, t] -= sum_i V_i[r, t] h[i ldh + t] (mg_maxpy with alpha = -1), then
 * out[i ldo + t] = sum_r conj(V_i[r, t]) w[r, t] with the updated w
 * (mg_mdot's order), for 1 <= k <= 16 (no reference counterpart: the two
 * middle passes of classical Gram-Schmidt twice in one read of V); work as
 * mg_mdot's. */
int mg_maxpy_mdot(const void* V, int64_t sV, int64_t ldv, int k, const void* h, int64_t ldh, void* w, int64_t ldw,
                  int64_t n, int64_t T, void* work, void* out, int64_t ldo, void* stream);
/* norms[t] = ||X[:, t]||_2 of an n x T complex128 matrix (row stride ldx),
 * summed in the fixed order of mg_maxpy_norm; work as there. */
int mg_colnorm(const void* X, int64_t ldx, int64_t n, int64_t T, void* work, double* norms, void* stream);
/* Relaxed-accuracy Krylov products of the refinement (no counterpart in the
 * reference): once the GMRES residual has dropped to r, a product with a
 * relative error of about tol / r moves the final residual by about tol, so
 * the last products of a solve may run on split-bf16 MFMA.
 * mg_split_rows: A (rows x cols, lda; prec MG_C128 or MG_C64) -> AS
 * (mg_split_rows_bytes bytes, 16-byte aligned), each element rounded to
 * complex float and held as three bf16 parts of its real and imaginary part.  mg_relaxed_product: C (m x n complex128, ldc) = A B with
 * A given as AS (m x k) and B (k x n complex128, ldb); prec MG_BF16X3 (three
 * bf16 products per real product, ~16 bits) or MG_BF16X6 (six, ~24 bits);
 * work: mg_relaxed_product_work_bytes(prec, m, n, k) bytes, 16-byte aligned.
 * k = 0 is MG_ERR_ARG. */
int64_t mg_split_rows_bytes(int64_t rows, int64_t cols);
int mg_split_rows(int prec, const void* A, int64_t lda, int64_t rows, int64_t cols, void* AS, void* stream);
int64_t mg_relaxed_product_work_bytes(int prec, int64_t m, int64_t n, int64_t k);
int mg_relaxed_product(int prec, int64_t m, int64_t n, int64_t k, const void* AS, const void* B, int64_t ldb,
                       void* C, int64_t ldc, void* work, int64_t work_bytes, void* stream);

/* Grouped LU: up to 256 independent matrices of uneven order factored in
 * lockstep, every launch of the blocked algorithm serving every problem with
 * work at that column, and their right-hand sides solved the same way.  It
 * replaces magreete's per-configuration loop (magreete.py:346-452), whose
 * every iteration factors one matrix in torch.linalg.solve
 * (Transmission3D.py:205, Transmission2D.py:192).  Per problem the factors,
 * pivots (0-based) and info (first zero pivot, 1-based) are getrf's, with
 * the same pivoting rule (MG_BF16X3 groups with tournament leaves, the
 * default: each leaf's pivots are those of partial pivoting among the rows
 * that chunks of 256 rows offer, ipiv keeping getrf's meaning).
 * mg_group_create: prec MG_C128, MG_C64 or MG_BF16X3 (complex float, trailing
 * updates on BF16x3 products); host arrays of `count` device pointers and
 * sizes (A[i]: n[i] x n[i], lda[i]; ipiv[i]: n[i] ints; info[i]: one int;
 * B[i]: n[i] x nrhs, ldb[i], or B = NULL and nrhs = 0).  The descriptors are
 * copied to a device table owned by the group; the buffers stay the
 * caller's and must outlive it.
 * mg_group_configure: panel width nb (multiple of 16, <= 1024), leaf width
 * (<= 16), trailing-update tile (0: 128 x 64, 1: 64 x 64), leaf workgroup
 * threads (256, 512, 1024), pivot (MG_BF16X3 leaves: 1 tournament, 0
 * partial pivoting as getrf; other precisions always pivot as getrf).  mg_group_tune: the same tunables chosen by
 * timing candidates on copies of a sample of the group (call before the
 * factorisation), cached per device, precision and size class; cfg (4 ints)
 * receives them, best_ms the sample's time (0 for a cached choice).
 * mg_group_getrf / mg_group_getrs: factor / solve on `stream`.
 * mg_group_capture(op 0: factor, 1: solve) records that launch sequence in
 * a HIP graph (nothing runs; no memset nodes) and mg_group_replay launches
 * it on `stream`: the buffers are the group's, their contents may change
 * between replays.
 * Calls on one group are not reentrant. */
typedef struct mg_group mg_group;
int mg_group_create(int prec, int count, void* const* A, const int64_t* n, const int64_t* lda, int* const* ipiv,
                    int* const* info, void* const* B, const int64_t* ldb, int64_t nrhs, mg_group** out);
int mg_group_destroy(mg_group* g);
int mg_group_configure(mg_group* g, int nb, int leaf, int tile, int leaf_threads, int pivot);
int mg_group_tune(mg_group* g, int* cfg, float* best_ms, void* stream);
int mg_group_getrf(mg_group* g, void* stream);
int mg_group_getrs(mg_group* g, void* stream);
int mg_group_capture(mg_group* g, int op);
int mg_group_replay(mg_group* g, int op, void* stream);
/* Iterative refinement over a group: mg_group_set_aux registers an n[i] x
 * nrhs block X[i] (ldx[i]) per problem (caller order); mg_group_residual
 * then forms B[i] -= A[i] X[i] for every problem in one launch (products in
 * the group's element type, so a complex128 group of the original matrices
 * gives the FP64 residual b - M x when B holds b); mg_group_colnorm writes
 * ||B[i][:, c]|| to norms[i nrhs + c] (fixed summation order).  No reference
 * counterpart: torch.linalg.solve (Transmission3D.py:205) refines nothing. */
int mg_group_set_aux(mg_group* g, void* const* X, const int64_t* ldx);
int mg_group_residual(mg_group* g, void* stream);
int mg_group_colnorm(mg_group* g, double* norms, void* stream);
/* Launches issued by the group's calls so far (measurement hook). */
int64_t mg_group_launches(const mg_group* g);

#ifdef __cplusplus
}
#endif
#endif
