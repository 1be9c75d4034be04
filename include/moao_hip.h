/* moao_hip.h -- C-ABI of the MI355X (gfx950) MOAO reconstructor path.
 *
 * Plain C: pointers, sizes and status codes only.  Every matrix is column-major with an
 * explicit leading dimension, exactly as in the reference.  Two flavours of each compute entry:
 *   - host-pointer entries  (moao_hip_<op>)      : drop-in for the reference's CPU functions;
 *     they stage through device memory internally and return when the result is back on the
 *     host, with the reference's in-place semantics;
 *   - device-pointer entries (moao_hip_<op>_dev) : operate on device buffers, stream-ordered on
 *     `stream` (NULL = the library stream).
 * Status: 0 = ok; >0 = LAPACK-style info (e.g. the order of the first non-positive pivot);
 * <0 = one of the MOAO_ERR_* codes below.  moao_last_error() returns a message for the
 * calling thread's last failure.
 */
#ifndef MOAO_HIP_H
#define MOAO_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOAO_OK 0
#define MOAO_ERR_ARG (-1)      /* invalid argument (the reference exits or misbehaves) */
#define MOAO_ERR_HIP (-2)      /* HIP runtime error */
#define MOAO_ERR_NOMEM (-3)    /* device allocation failed / over the HBM budget */
#define MOAO_ERR_NOCONV (-4)   /* iterative refinement did not reach its tolerance */
#define MOAO_ERR_IO (-5)       /* parameter file missing, truncated or inconsistent */
#define MOAO_ERR_NODEV (-6)    /* no usable gfx950 device */

/* Precision modes of the reconstructor. */
#define MOAO_FP64 0            /* FP64 everywhere: the reference's lapack arithmetic */
#define MOAO_MIXED_FP32 1      /* FP32 trailing updates, FP64 diagonal tiles + FP64 refinement */
#define MOAO_MIXED_BF16 2      /* BF16 MFMA trailing updates (each FP32 operand split into two
                                  BF16 terms, FP32 accumulation), FP64 diagonal tiles + FP64
                                  refinement */

const char* moao_last_error(void);
const char* moao_version(void);

/* ---------------------------------------------------------------- system model (host) ----
 * Replaces Tomo_struct<double> (common/tomo_struct/tomo_struct.hpp:9-94): reads the
 * reference's sys-params / atm-params text files and prepares the vectors matcov reads.   */
typedef struct moao_tomo moao_tomo;

/* POD view of what the matcov generator reads; mirrors the argument list of the reference's
 * matcov_tile_gpu_4 (common/kernels/matcov_kernels_gpu.hpp:15-20) plus the truth-sensor
 * X, Y used by Ctt (matcov_ts_kernel_tile, common/kernels/matcov_kernels.cpp:495-515).
 * All pointers are HOST pointers.  Callers holding their own Tomo_struct fill it from its
 * fields (see moao_hip.hpp). */
typedef struct moao_tomo_view {
    int Nw;                 /* number of WFS including the truth sensor (last) */
    int Nlayer;             /* turbulent layers */
    int nlgs;               /* LGS count (the first nlgs WFS) */
    long Nx;                /* valid subapertures over all WFS */
    double diam;            /* telescope diameter (m) */
    double lgs_cst;         /* constant added to LGS xx/yy blocks */
    const long* Nsubap;     /* [Nw] */
    const long* Nssp;       /* [Nw] */
    const double* sspSizeL; /* [Nw*Nlayer], w-major */
    const double* u;        /* [Nlayer*Nx], layer-major */
    const double* v;        /* [Nlayer*Nx] */
    const double* X;        /* [Nx] unprojected subaperture centres */
    const double* Y;        /* [Nx] */
    const double* L0;       /* [Nlayer] outer scale per layer (the reference's L0diff) */
    const double* cn2;      /* [Nlayer] */
    const double* noiseNGS; /* [Nw-nlgs] */
    const double* noiseLGSxx; /* [Nx] */
    const double* noiseLGSyy; /* [Nx] */
    const double* noiseLGSxy; /* [Nx] */
} moao_tomo_view;

/* Tomo_struct(sys_file, atm_file)  (tomo_struct.cpp:11-17) */
int moao_tomo_create(const char* sys_file, const char* atm_file, moao_tomo** out);
/* Tomo_struct::updateAtm(atm_file)  (tomo_struct.cpp:212-262) */
int moao_tomo_update_atm(moao_tomo* t, const char* atm_file);
/* Tomo_struct::setTarget(i)  (tomo_struct.cpp:275-279) */
int moao_tomo_set_target(moao_tomo* t, int target);
/* Tomo_struct::setTargetCoord(aX, aY, toRad)  (tomo_struct.cpp:291-303) */
int moao_tomo_set_target_coord(moao_tomo* t, double ax, double ay, int to_rad);
/* Fills *view with pointers into t (valid until the next update/set call or destroy). */
int moao_tomo_get_view(const moao_tomo* t, moao_tomo_view* view);
/* getNMeas / getNMeasTS  (tomo_struct.cpp:569-579) */
long moao_tomo_nmeas(const moao_tomo* t);
long moao_tomo_nmeasts(const moao_tomo* t);
/* nW, nLgs, nNgs, nTarget, nLayer, Nssp: a small int summary (out[6]) */
int moao_tomo_dims(const moao_tomo* t, long* out6);
/* Copies a named vector field ("X","Y","u","v","sspSizeL","L0","cn2","noiseNGS",
 * "noiseLGSxx","noiseLGSyy","noiseLGSxy","lgsExt","lgsTheta","Nsubap","Nssp") into out
 * (as double); returns its length, or MOAO_ERR_ARG if the name is unknown.  out may be NULL
 * to query the length. */
long moao_tomo_field(const moao_tomo* t, const char* name, double* out, long cap);
void moao_tomo_destroy(moao_tomo* t);

/* ---------------------------------------------------------------- device / memory --------- */
/* Selects and initialises the device (one process, one GPU).  Called implicitly with device 0
 * by the first compute entry if the caller did not. */
int moao_hip_init(int device);
int moao_hip_finalize(void);
int moao_hip_malloc(void** dptr, size_t bytes);
int moao_hip_free(void* dptr);
int moao_hip_memcpy_h2d(void* dst, const void* src, size_t bytes);
int moao_hip_memcpy_d2h(void* dst, const void* src, size_t bytes);
int moao_hip_memcpy_d2d(void* dst, const void* src, size_t bytes);
int moao_hip_synchronize(void);
/* Device memory in use / total, bytes. */
int moao_hip_mem_info(size_t* free_bytes, size_t* total_bytes);
/* Streams for the stream-ordered entries (default priority; NULL means the library stream).
 * No reference counterpart: the reference's CPU calls are synchronous. */
int moao_hip_stream_create(void** stream);
int moao_hip_stream_synchronize(void* stream);
int moao_hip_stream_destroy(void* stream);

/* ---------------------------------------------------------------- matcov ------------------
 * MOAO_COMMON::matcov_tile (common/compute/matcov.hpp:14-18, matcov.cpp:17-98).
 * type_mat: 1 = Cmm (only global row >= global col is written), 2 = Cmaa (full, reference
 * quirk kept), 3 = Ctm (full), 4 = Ctt (full).  xoffset is the global row offset and yoffset
 * the global column offset of the tile; element (lx, ly) goes to data[ly*lda + lx].
 * Unknown type_mat returns MOAO_ERR_ARG (the reference calls exit(1)).                    */
int moao_hip_matcov_tile(double* data, int nrows, int ncols, int xoffset, int yoffset,
                         int lda, const moao_tomo_view* tomo, int type_mat);
int moao_hip_matcov_tile_dev(double* d_data, int nrows, int ncols, int xoffset, int yoffset,
                             int lda, const moao_tomo_view* tomo, int type_mat, void* stream);

/* ---------------------------------------------------------------- reconstructor -----------
 * MOAO_LAPACK::reconstructor (lapack/compute/reconstructor.hpp:19-21, reconstructor.cpp:32-58):
 * dpotrf('L', Cmm) then, per target g, Ctm_g = Ctm + g*nmeas*nmeasts is overwritten by
 * R_g = Ctm_g * Cmm^{-1} (dtrsm Right/Lower/Trans then Right/Lower/NoTrans), FP64.
 * Returns 0, the dpotrf info (> 0: leading minor of that order not positive definite; Cmm is
 * left partially factored -- complete in the columns of the 128-column leaves before the
 * failing pivot's leaf, and in that leaf's own rows of its 16-column blocks before the failing
 * one -- and, for the host entry, Ctm untouched), or MOAO_ERR_*.            */
int moao_hip_reconstructor(long nmeas, long nmeasts, double* Cmm, long ldCmm, double* Ctm,
                           long ldCtm, long ngal);
/* Device buffers; returns when the solve is done (its status is read back).  On info > 0 the
 * Ctm blocks hold partial results.  Workspace comes from the stream-ordered pool, per call, so
 * calls on different streams may overlap. */
int moao_hip_reconstructor_dev(long nmeas, long nmeasts, double* dCmm, long ldCmm, double* dCtm,
                               long ldCtm, long ngal, void* stream);
/* The same solve fully stream-ordered: returns once the work is queued on `stream`; the dpotrf
 * status is written to the device int *d_info when the solve runs (0, or the order of the first
 * non-positive pivot, after which every later kernel of the solve does nothing).  No
 * counterpart in the reference (its calls are synchronous); for callers that pipeline solves. */
int moao_hip_reconstructor_async(long nmeas, long nmeasts, double* dCmm, long ldCmm,
                                 double* dCtm, long ldCtm, long ngal, int* d_info, void* stream);

/* The same solve as moao_hip_reconstructor_async with no allocation of any kind: its workspace
 * (the inverse leaves) is a static device array of 40 MiB that each device holds for the life
 * of the process, so nmeas is limited to 40960 (larger orders return MOAO_ERR_ARG).  *d_info
 * is used as the solve's status word and holds the dpotrf status when the solve has run.  No
 * counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, fixed serving loops).  Every call on one device shares the array, so
 * calls that overlap on one device (different streams, or threads) race; calls ordered on one
 * stream do not.  Call moao_hip_init first when the caller must not create anything: the
 * library's init takes the scratch address (and creates the library stream used when stream
 * is NULL). */
int moao_hip_reconstructor_noalloc(long nmeas, long nmeasts, double* dCmm, long ldCmm,
                                   double* dCtm, long ldCtm, long ngal, int* d_info,
                                   void* stream);

/* A captured FP64 solve: the launch sequence of moao_hip_reconstructor_async on these buffers
 * recorded once as a HIP graph (no allocation, memset or host synchronisation inside; the leaf
 * inverses live in workspace the graph object owns) and replayed by solve_graph_launch on any
 * stream.  Each launch solves whatever Cmm / Ctm hold at that point (refill them in between, as
 * the reference's refine loop regenerates them); the status goes to *d_info.  No counterpart
 * in the reference: for callers that solve one shape repeatedly, where launch work is a large
 * part of a small solve.  A graph must not run on two streams at once (shared workspace). */
typedef struct moao_solve_graph moao_solve_graph;
int moao_hip_solve_graph_create(long nmeas, long nmeasts, double* dCmm, long ldCmm, double* dCtm,
                                long ldCtm, long ngal, int* d_info, moao_solve_graph** out);
int moao_hip_solve_graph_launch(moao_solve_graph* g, void* stream);
void moao_hip_solve_graph_destroy(moao_solve_graph* g);

/* The same solve on FP32 storage: MOAO_CHAMELEON::reconstructor<float>, the precision the
 * reference's Chameleon driver runs in production (chameleon/moao_chameleon.cpp:41,
 * chameleon/compute/reconstructor.cpp:31-65: spotrf + 2 strsm).  FP32 MFMA updates, the
 * diagonal 128 x 128 blocks factored in FP64 and rounded to FP32.  Status as the FP64 entry
 * (> 0: the FP32 factorization met a non-positive pivot, which FP32 does on ill-conditioned
 * Cmm, as spotrf would). */
int moao_hip_sreconstructor(long nmeas, long nmeasts, float* Cmm, long ldCmm, float* Ctm,
                            long ldCtm, long ngal);
int moao_hip_sreconstructor_dev(long nmeas, long nmeasts, float* dCmm, long ldCmm, float* dCtm,
                                long ldCtm, long ngal, void* stream);

/* Tile descriptors and sequences: MOAO_CHAMELEON::reconstructor<T>(Cmm, Ctm, sequence, check,
 * sync) (chameleon/compute/reconstructor.hpp:15-20, reconstructor.cpp:31-65).  A descriptor is
 * a device matrix in tile layout: tile (i, j) of mb x nb elements (the reference's ts),
 * column-major with leading dimension mb, at element offset (i + j*mt)*mb*nb with
 * mt = ceil(lm/mb), edge tiles padded to mb x nb (MORSE_Lapack_to_Tile's role is
 * moao_hip_lapack_to_tile).  A sequence is a stream plus a status: calls return once queued,
 * moao_hip_sequence_wait (MORSE_Sequence_Wait) blocks and returns the first non-zero status
 * (a dpotrf info > 0 or MOAO_ERR_*).  reconstructor_tile_async: Cmm (n x n) -> its factor,
 * Ctm (m x n) -> R, FP64 (dtype 0) or FP32 storage (dtype 1, the Chameleon production
 * precision); sync != 0 waits after the solve; check is ignored (a TODO in the reference). */
typedef struct moao_hip_desc {
    int dtype;     /* 0 = double, 1 = float */
    void* mat;     /* device memory in tile layout */
    int mb, nb;    /* tile size */
    long lm, ln;   /* matrix rows, columns */
} moao_hip_desc;
typedef struct moao_hip_sequence moao_hip_sequence;
int moao_hip_sequence_create(moao_hip_sequence** seq);
int moao_hip_sequence_wait(moao_hip_sequence* seq);
void moao_hip_sequence_destroy(moao_hip_sequence* seq);
int moao_hip_lapack_to_tile(const void* dA, long lda, moao_hip_desc* desc,
                            moao_hip_sequence* seq);
int moao_hip_tile_to_lapack(const moao_hip_desc* desc, void* dA, long lda,
                            moao_hip_sequence* seq);
int moao_hip_reconstructor_tile_async(const moao_hip_desc* Cmm, moao_hip_desc* Ctm,
                                      moao_hip_sequence* seq, int check, int sync);

/* Mixed precision (no reference counterpart; same result contract as the FP64 entry to a
 * stated tolerance).  Cmm is read only (lower triangle) and preserved; the Ctm block (one
 * target) is overwritten by R.  If `lowrank` describes the system Cmm was generated from, the
 * exact lgs_cst low-rank term of the LGS blocks is removed before the FP32 factorization and
 * restored by the Woodbury identity in FP64.  Refinement stops when
 * ||Ctm - R Cmm||_F <= tol ||Ctm||_F (returned in *resid) or after max_iter corrections
 * (then MOAO_ERR_NOCONV with R holding the last iterate).  mode: MOAO_MIXED_FP32 or
 * MOAO_MIXED_BF16.                                                                          */
int moao_hip_reconstructor_mixed_dev(long nmeas, long nmeasts, const double* dCmm, long ldCmm,
                                     double* dCtm, long ldCtm, const moao_tomo_view* lowrank,
                                     int mode, double tol, int max_iter, double* resid,
                                     int* iters, void* stream);

/* ---------------------------------------------------------------- Cee / Cvv ---------------
 * MOAO_LAPACK::Cee_Cvv (lapack/compute/cee_cvv.hpp:21, cee_cvv.cpp:50-124), FP64:
 *   Cee = Ctt - Ctm R^T - R Ctm^T + R Cmm R^T  (m x m),  Tmp = Dx Cee,  Cvv = Tmp Dx^T.
 * Cmm: regenerated covariance (lower triangle read).  Cpp (= Ctt) and Cpm (= Ctm) are
 * clobbered as in the reference: Cpp's lower part holds Ctt - Ctm R^T - R Ctm^T and Cpm holds
 * R Cmm on return.  Every ld argument is honoured.                                          */
int moao_hip_cee_cvv(long nmeas, long nmeasts, long nact, double* Cmm, long ldCmm, double* Cpp,
                     long ldCpp, double* Cpm, long ldCpm, double* R, long ldR, double* Dx,
                     long ldDx, double* Cee, long ldCee, double* Cvv, long ldCvv, double* Tmp,
                     long ldTmp);
int moao_hip_cee_cvv_dev(long nmeas, long nmeasts, long nact, const double* dCmm, long ldCmm,
                         double* dCpp, long ldCpp, double* dCpm, long ldCpm, const double* dR,
                         long ldR, const double* dDx, long ldDx, double* dCee, long ldCee,
                         double* dCvv, long ldCvv, double* dTmp, long ldTmp, void* stream);
/* Cee/Cvv of ngal targets in one call: the observation step's per-target Cee_Cvv calls
 * (lapack/moao_lapack.cpp:165-175) batched.  The targets share Cmm and Dx; target g's Cpp, Cpm,
 * R, Cee, Cvv and Tmp blocks start g block strides in (g*ldCpp*nmeasts, g*ldCpm*nmeas,
 * g*ldR*nmeas, g*ldCee*nmeasts, g*ldCvv*nact, g*ldTmp*nmeasts doubles), as the
 * reconstructor's targets do.  Every output equals that of ngal moao_hip_cee_cvv_dev calls bit
 * for bit; up to 8 targets' products share each launch.                                      */
int moao_hip_cee_cvv_batched_dev(long nmeas, long nmeasts, long nact, long ngal,
                                 const double* dCmm, long ldCmm, double* dCpp, long ldCpp,
                                 double* dCpm, long ldCpm, const double* dR, long ldR,
                                 const double* dDx, long ldDx, double* dCee, long ldCee,
                                 double* dCvv, long ldCvv, double* dTmp, long ldTmp,
                                 void* stream);

/* ---------------------------------------------------------------- intersample PSF ---------
 * intersample_prepare/init/process/free (common/tomo_struct/intersample.hpp:46-63,
 * intersample.cpp:72-333).  The reference reads idx/abs2fi/otftel from FITS files; here the
 * caller passes the arrays (idx: nidx 1-based map indices; abs2fi, otf_tel: N*N in the FITS /
 * reference flat layout).  N <= 0 selects the reference's rule (moao_intersample_default_n).
 * precision: 0 = FP32 FFTs (default product path), 1 = FP64 FFTs.
 * process: Cvv is the flat nact x nact covariance (nidx doubles), the PSF (N*N doubles, the
 * reference's `dphi` output layout) is written to psf; the _dev entry takes `batch` stacked
 * Cvv / PSF arrays (one per direction) in device memory.                                   */
typedef struct moao_isample moao_isample;
long moao_intersample_default_n(long np);
int moao_hip_intersample_create(long nidx, long np, long N, const long* idx, const double* abs2fi,
                                const double* otf_tel, int precision, moao_isample** out);
long moao_hip_intersample_n(const moao_isample* s);
int moao_hip_intersample_process(moao_isample* s, const double* Cvv, double* psf);
int moao_hip_intersample_process_dev(moao_isample* s, const double* dCvv, int batch,
                                     double* dPsf, void* stream);
void moao_hip_intersample_destroy(moao_isample* s);
/* Long-exposure PSF: dSum[i] += dPsf[i] for i < count (device memory, stream-ordered).  The
 * sum the reference's TODO describes (chameleon/moao_chameleon.cpp:324-345: the cumulated PSF
 * of each direction over the run), kept on the device so an observation loop that writes no
 * short-exposure PSF never copies one to the host.                                          */
int moao_hip_psf_accumulate_dev(const double* dPsf, long count, double* dSum, void* stream);

/* The reference's own intersample boundary, same names, arguments and results
 * (common/tomo_struct/intersample.hpp:22-63, intersample.cpp:175-333):
 *   prepare  N = max(512, 8*2^(long)(ln(np)+0.5)), dactupix = N/(3(np-1))+1 (long division),
 *            lambda = 1.65e-6, file names path/idx.fits, path/abs2fi.fits, path/otftel.fits;
 *            Dtel is ignored as in the reference;
 *   init     reads the three FITS files (first axis checked against nidx / N, as the
 *            reference does) and sets up the device stage;
 *   process  the PSF of Cvv (nidx values, double or float) is left in dphi (N*N doubles);
 *   free     releases the arrays and the device stage.
 * The struct keeps the reference's public fields; FFTW's plans and scratch are replaced by the
 * device stage `dev`.  precision (set by prepare to 1 = FP64 transforms, the reference's
 * arithmetic) may be set to 0 before init for the FP32-transform product path.
 * Returns MOAO_* status codes (the C++ layer in moao_hip.hpp maps them to the reference's
 * 1 = success / 0 = failure). */
typedef struct moao_isample_struct {
    long N;            /* size of the FFT support */
    long np;           /* actuators across the diameter (Nssp of the truth sensor + 1) */
    long nidx;         /* entries of the index map (nact * nact) */
    double lambda;     /* wavelength */
    double dactupix;   /* inter-actuator distance in pixels */
    long* idx;         /* [nidx] 1-based map indices (read by init) */
    double* abs2fi;    /* [N*N] */
    double* otf_tel;   /* [N*N] */
    double* dphi;      /* [N*N] the PSF after process */
    char filename_idx[256], filename_abs2fi[256], filename_otftel[256];
    int precision;     /* 1 = FP64 transforms (default), 0 = FP32 transforms */
    moao_isample* dev; /* device stage (init) */
} moao_isample_struct;
void moao_hip_intersample_prepare(moao_isample_struct* s, long nidx, long np, float Dtel,
                                  const char* path);
int moao_hip_intersample_init(moao_isample_struct* s);
int moao_hip_intersample_process_d(moao_isample_struct* s, const double* Cvv);
int moao_hip_intersample_process_f(moao_isample_struct* s, const float* Cvv);
void moao_hip_intersample_free(moao_isample_struct* s);

/* ---------------------------------------------------------------- accuracy check ----------
 * compareMatrices (common/compute/check_utils.cpp:20-57), the reference check harness's metric:
 * max_err = max |1 - B/A| over the visited entries (|B| where A == 0), frob = sqrt of the sum of
 * (B - A)^2 (off-diagonal terms doubled for uplo 'U' / 'L').  A is column-major nrows x ncols;
 * B's entry (i, j) is B[i + j*nrows] when trans == 'T', else B[i*ncols + j].  uplo 'U' visits
 * j >= i, 'L' j < i, anything else every entry.  The _dev variant reads device buffers (one
 * reduction launch) and returns when the result is on the host.                            */
int moao_compare_matrices(const double* A, const double* B, long nrows, long ncols, char uplo,
                          char trans, double* max_err, double* frob);
int moao_hip_compare_matrices_dev(const double* dA, const double* dB, long nrows, long ncols,
                                  char uplo, char trans, double* max_err, double* frob,
                                  void* stream);

/* ---------------------------------------------------------------- FITS images -------------
 * The reference's FITS layout (common/io/fits.cpp:66-148): a primary image, NAXIS1 the fast
 * axis, memory order on disk.  read: any of BITPIX -64/-32/8/16/32/64 as doubles; out == NULL
 * queries the axes only (cap = capacity of out, in values).  write: doubles (writeFits). */
int moao_fits_read(const char* path, long* nx, long* ny, double* out, long cap);
int moao_fits_write(const char* path, long nx, long ny, const double* data);

/* ---------------------------------------------------------------- GEMM utility ------------
 * C = alpha*A*op(B) + beta*C on the MFMA engine the stages use (column-major);
 * op(B) = B^T when transb == 1 (B is N x K), B when transb == 0 (B is K x N), and for
 * transb == 2 the symmetric K x K matrix whose lower triangle B holds.  No reference
 * counterpart: exposed for tests and for callers composing their own chains.                */
int moao_hip_dgemm_dev(int transb, int M, int N, int K, double alpha, const double* A, long lda,
                       const double* B, long ldb, double beta, double* C, long ldc, void* stream);
int moao_hip_sgemm_dev(int transb, int M, int N, int K, float alpha, const float* A, long lda,
                       const float* B, long ldb, float beta, float* C, long ldc, void* stream);
/* The FP32 product on BF16 MFMA: each operand element split into two BF16 terms (hi =
 * bf16(a), lo = bf16(a - hi)), hi*hi + hi*lo + lo*hi accumulated in FP32 (about 2^-16 relative
 * per product).  transb 0 or 1.  The kernel of the mixed BF16 reconstructor's updates. */
int moao_hip_sgemm_bf16x3_dev(int transb, int M, int N, int K, float alpha, const float* A,
                              long lda, const float* B, long ldb, float beta, float* C, long ldc,
                              void* stream);

/* ---------------------------------------------------------------- timing ------------------
 * Per-kernel-class timing with HIP events recorded on each launch's stream.  No reference
 * counterpart (the reference times with gettimeofday around whole calls).  Classes:
 * 0 FP64 GEMM, 1 FP32 GEMM, 2 leaf factor, 3 matcov, 4 precision conversion, 5 FFT, 6 other.
 * work[] is the algorithmic work of the class: flops for GEMM/leaf, elements written for
 * matcov, bytes moved for conversion/FFT passes.                                           */
int moao_hip_profile_begin(void);
int moao_hip_profile_end(double* ms, double* work, long* count);  /* arrays of 7 */

/* ---------------------------------------------------------------- serving -----------------
 * Continuous batching of reconstructor requests (new systems / atmospheric profiles) on one
 * GPU.  The reference drives one system at a time from its host loop
 * (lapack/moao_lapack.cpp:95-188: Tomo from files -> matcov Cmm, Ctm -> reconstructor); a
 * request here is that refine step for one (sys-params, atm-params, target).  submit() reads
 * the files and sizes the request (n, m, HBM footprint) on the caller's thread and returns at
 * once; `slots` device streams each take the next request that fits the byte budget as soon as
 * they are free, so requests arriving while others run are admitted as slots and memory free
 * up, and their kernels overlap.  A request's kernels and inputs do not depend on the mix: its
 * R is bit-identical alone and in a mix.                                                    */
typedef struct moao_server moao_server;

typedef struct moao_serve_config {
    int slots;          /* concurrent requests on the device (0 -> 4, at most 16) */
    size_t hbm_budget;  /* bytes of device memory requests may hold together (0 -> 85% free) */
    int max_bypass;     /* times a queued request may be overtaken by smaller ones (0 -> 8) */
} moao_serve_config;

typedef struct moao_serve_request {
    const char* sys_file;  /* reference-format sys-params file */
    const char* atm_file;  /* reference-format atmospheric profile */
    int target;            /* truth direction (Tomo_struct::setTarget) */
    int mode;              /* MOAO_FP64, MOAO_MIXED_FP32 or MOAO_MIXED_BF16 */
    double tol;            /* mixed: relative residual to refine to (0 -> 1e-11) */
    int max_iter;          /* mixed: refinement steps allowed (0 -> 5) */
    double* R;             /* host output m x n (column-major) or NULL; written before wait */
    long ldR;              /* leading dimension of R (>= m) */
    long R_cols;           /* columns R holds; must equal n (checked at submit) */
} moao_serve_request;

typedef struct moao_serve_result {
    int status;            /* as the compute entries: 0, >0 dpotrf info, <0 MOAO_ERR_* */
    int slot;
    long n, m;
    int iters;             /* mixed: refinement steps taken */
    double resid;          /* mixed: final relative residual */
    size_t footprint;      /* bytes of device memory the request held */
    double t_submit, t_start, t_end;  /* seconds since the server was created */
} moao_serve_result;

typedef struct moao_serve_counters {
    long queued, running, completed, peak_running;
    size_t bytes_in_use, peak_bytes, budget;
    long graph_captures;   /* FP64 solves recorded as graphs (a slot's solve seen twice) */
    long graph_replays;    /* FP64 solves run by replaying a slot's graph */
} moao_serve_counters;

int moao_serve_create(const moao_serve_config* cfg, moao_server** out);
/* MOAO_ERR_NOMEM when the request alone exceeds the budget; MOAO_ERR_IO on unreadable files. */
int moao_serve_submit(moao_server* s, const moao_serve_request* req, long* ticket);
/* Blocks until the request is done; returns its status and forgets the ticket. */
int moao_serve_wait(moao_server* s, long ticket, moao_serve_result* out);
int moao_serve_stats(moao_server* s, moao_serve_counters* out);
/* Finishes every submitted request, then stops the slots. */
void moao_serve_destroy(moao_server* s);

#ifdef __cplusplus
}
#endif

#endif /* MOAO_HIP_H */
