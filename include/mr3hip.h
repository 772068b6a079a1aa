/*
 * mr3hip.h -- public C ABI of the MI355X MRRR tridiagonal eigensolver (libmr3hip.so).
 *
 * Plain pointers and sizes only; every entry point is extern "C".
 */
#ifndef MR3HIP_H
#define MR3HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Drop-in replacement of the reference routine
 *   int mrrr(char *jobz, char *range, int *n, double *D, double *E, double *vl,
 *            double *vu, int *il, int *iu, int *tryrac, int *m, double *W,
 *            double *Z, int *ldz, int *Zsupp);
 * (HPAC/mr3smp INCLUDE/mrrr.h:62-65, SRC/mrrr.c:87-296).  Same argument meaning:
 * jobz "N"/"V"/"C", range "A"/"V"/"I" (first character, any case), host arrays,
 * Z column-major with leading dimension ldz, Zsupp 1-based supports.  Returns 0 on
 * success, 1 for an invalid argument (mrrr.c:118-126), 2 when the device or the
 * algorithm fails (the reference only asserts there).  On return D and E hold no
 * defined values, as in the reference.  Safe to call from several host threads.
 */
int mrrr(char *jobz, char *range, int *n, double *D, double *E, double *vl,
         double *vu, int *il, int *iu, int *tryrac, int *m, double *W, double *Z,
         int *ldz, int *Zsupp);

/* Fortran binding of the same routine (SRC/mrrr.c:630-637): info receives the
 * return value. */
void mrrr_(char *jobz, char *range, int *n, double *D, double *E, double *vl,
           double *vu, int *il, int *iu, int *tryrac, int *m, double *W, double *Z,
           int *ldz, int *Zsupp, int *info);

/*
 * Dense drivers.  Each replaces the reference routine of the same name and prototype
 * (INCLUDE/mrrr.h:270-806, SRC/dense.c:151-1114), with the same argument meaning:
 * jobz "N"/"V"/"C", range "A"/"V"/"I", uplo "L"/"U" (first character, any case),
 * column-major host arrays.  The matrix is scaled into the safe range as
 * dense.c:476-566 and :940-1034 do, reduced to tridiagonal form on the GPU
 * (rocSOLVER), solved by the tridiagonal MRRR pipeline and back-transformed on the GPU.
 * Packed arrays (AP, BP) hold the uplo triangle by columns.  Complex arrays are
 * interleaved (re, im) pairs of doubles -- the layout of C99 double complex,
 * std::complex<double> and Fortran COMPLEX*16 -- passed as void *.
 *
 * Standard problem (dsyeig, dspeig, zheeig, zhpeig): W gets m eigenvalues, Z
 * (ldz x m) the eigenvectors.  A / AP is not modified: the reference leaves the
 * Householder vectors of its reduction there, which no caller can use since their
 * scalar factors are freed.
 * Generalized problem (dsygeig, zhegeig: itype 1 A x = lambda B x, 2 A B x =
 * lambda x, 3 B A x = lambda x): the first m columns of A receive the eigenvectors
 * and B its Cholesky factor, as in the reference; the packed forms (dspgeig,
 * zhpgeig) return the eigenvectors in Z and the packed factor in BP.
 *
 * Returns 0; 1 for an invalid argument (the checks of dense.c:174-180, :256-260,
 * :341-347, :422-426 and their complex twins, plus a leading dimension below n and, in the
 * generalized drivers, the range checks of the standard ones); 2 if B is not positive
 * definite or the device fails (the reference asserts there).
 */
int dsyeig(char *jobz, char *range, char *uplo, int *n, double *A, int *lda,
           double *vl, double *vu, int *il, int *iu, int *m, double *W, double *Z,
           int *ldz);                                        /* mrrr.h:270-272 */
int dspeig(char *jobz, char *range, char *uplo, int *n, double *AP, double *vl,
           double *vu, int *il, int *iu, int *m, double *W, double *Z,
           int *ldz);                                        /* mrrr.h:444-446 */
int dsygeig(int *itype, char *jobz, char *range, char *uplo, int *n, double *A,
            int *lda, double *B, int *ldb, double *vl, double *vu, int *il, int *iu,
            int *m, double *W);                              /* mrrr.h:352-356 */
int dspgeig(int *itype, char *jobz, char *range, char *uplo, int *n, double *AP,
            double *BP, double *vl, double *vu, int *il, int *iu, int *m,
            double *W, double *Z, int *ldz);                 /* mrrr.h:526-530 */
int zheeig(char *jobz, char *range, char *uplo, int *n, void *A, int *lda,
           double *vl, double *vu, int *il, int *iu, int *m, double *W, void *Z,
           int *ldz);                                        /* mrrr.h:538-541 */
int zhpeig(char *jobz, char *range, char *uplo, int *n, void *AP, double *vl,
           double *vu, int *il, int *iu, int *m, double *W, void *Z,
           int *ldz);                                        /* mrrr.h:718-721 */
int zhegeig(int *itype, char *jobz, char *range, char *uplo, int *n, void *A,
            int *lda, void *B, int *ldb, double *vl, double *vu, int *il, int *iu,
            int *m, double *W);                              /* mrrr.h:624-628 */
int zhpgeig(int *itype, char *jobz, char *range, char *uplo, int *n, void *AP,
            void *BP, double *vl, double *vu, int *il, int *iu, int *m, double *W,
            void *Z, int *ldz);                              /* mrrr.h:802-806 */

/* Fortran bindings of the dense drivers (SRC/dense.c:1037-1114): info receives the
 * return value. */
void dsyeig_(char *jobz, char *range, char *uplo, int *n, double *A, int *lda,
             double *vl, double *vu, int *il, int *iu, int *m, double *W, double *Z,
             int *ldz, int *info);
void dspeig_(char *jobz, char *range, char *uplo, int *n, double *AP, double *vl,
             double *vu, int *il, int *iu, int *m, double *W, double *Z, int *ldz,
             int *info);
void dsygeig_(int *itype, char *jobz, char *range, char *uplo, int *n, double *A,
              int *lda, double *B, int *ldb, double *vl, double *vu, int *il, int *iu,
              int *m, double *W, int *info);
void dspgeig_(int *itype, char *jobz, char *range, char *uplo, int *n, double *AP,
              double *BP, double *vl, double *vu, int *il, int *iu, int *m,
              double *W, double *Z, int *ldz, int *info);
void zheeig_(char *jobz, char *range, char *uplo, int *n, void *A, int *lda,
             double *vl, double *vu, int *il, int *iu, int *m, double *W, void *Z,
             int *ldz, int *info);
void zhpeig_(char *jobz, char *range, char *uplo, int *n, void *AP, double *vl,
             double *vu, int *il, int *iu, int *m, double *W, void *Z, int *ldz,
             int *info);
void zhegeig_(int *itype, char *jobz, char *range, char *uplo, int *n, void *A,
              int *lda, void *B, int *ldb, double *vl, double *vu, int *il, int *iu,
              int *m, double *W, int *info);
void zhpgeig_(int *itype, char *jobz, char *range, char *uplo, int *n, void *AP,
              void *BP, double *vl, double *vu, int *il, int *iu, int *m, double *W,
              void *Z, int *ldz, int *info);

/*
 * dsyeig on DEVICE arrays: dA (lda x n) is overwritten by the reduction, dW gets m
 * eigenvalues and dZ (ldz x zcols) the eigenvectors, on the given stream (NULL = the
 * library's per-thread stream).  vl, vu, il, iu and m are host scalars as in dsyeig.
 * No counterpart in the reference, whose dense routines take host arrays only.
 */
int mr3_dsyeig_device(char jobz, char range, char uplo, int n, double *dA, int lda,
                      double *vl, double *vu, int *il, int *iu, int *m, double *dW,
                      double *dZ, int ldz, int zcols, void *stream);

/*
 * One problem of a batched solve.  All array pointers are DEVICE pointers on the
 * current HIP device; D and E are read only.  Inputs: n, jobz ('N', 'V' or 'C'),
 * range ('A', 'V' or 'I'), vl/vu or il/iu as for mrrr(), tryrac, ldz and mcols
 * (columns available in Z).  Outputs: W (n entries), Z (ldz x m, column-major),
 * Zsupp (2m ints), and the scalar fields m, info, vl_out, vu_out, il_out,
 * iu_out, tryrac_out.  Counterpart of one mrrr() call (mrrr.h:62-65); it has no
 * batched twin in the reference, whose callers loop over mrrr().
 */
typedef struct mr3_problem {
    int n;
    char jobz, range;
    char pad0, pad1;
    double vl, vu;
    int il, iu;
    int tryrac;
    int ldz, mcols;
    const double *D, *E;
    double *W, *Z;
    int *Zsupp;
    /* outputs */
    int m, info;
    double vl_out, vu_out;
    int il_out, iu_out, tryrac_out;
} mr3_problem;

/*
 * Solves nprob independent problems in one pass of the device pipeline on the
 * given HIP stream (NULL = the library's per-thread stream).  Returns 0 if every
 * problem succeeded, 1 if any had invalid arguments, 2 on a device failure; the
 * per-problem result is in probs[i].info.  The call returns when the results are
 * in device memory and the scalar outputs are filled in.
 */
int mr3_solve_batch(mr3_problem *probs, int nprob, void *stream);

/* The single-problem device path: mr3_solve_batch with nprob = 1. */
int mr3_solve_one(mr3_problem *prob, void *stream);

/*
 * mr3_solve_batch for problems of order n < 500 (the reference's dstemr route,
 * SRC/mrrr.c:149-153) that allocates nothing, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The work arrays live in a static device buffer of
 * 64 MiB that each GPU the library runs on holds until the process ends, and the
 * descriptors are staged in static host arrays; the caller passes the inputs and
 * outputs as for mr3_solve_batch (stream NULL: the calling thread's default stream).
 * Same results, bit for bit, and return codes as mr3_solve_batch, plus 3 -- nothing
 * run, every info 3 -- when a problem has n >= 500, nprob exceeds 1024, or the
 * problems' work arrays exceed the buffer (about 1.6 MB for n = 499, proportional to
 * n below).  The calls of a process are serialised; calls that overlap on one GPU
 * from several processes share the buffer and race.
 */
int mr3_solve_small_noalloc(mr3_problem *probs, int nprob, void *stream);

/*
 * A batched solve captured as one HIP graph, for a caller that solves the same
 * problems (the same device arrays; D and E may change between launches) again and
 * again: the small-matrix route and the task-parallel pipeline with a fixed number of
 * tree levels and one pool for the child representations, in device memory the graph
 * owns, launched with no host work inside.  No counterpart in the reference (the
 * reference walks its representation tree with a pthreads work queue, mrrr_vec.c:
 * 226-299, queue.c); this is the replayable form of mr3_solve_batch.
 * mr3_graph_create solves the batch twice as mr3_solve_batch does (to learn the depth
 * of its tree and the slots it needs; the problems' outputs are valid afterwards) and
 * captures it; it returns what mr3_solve_batch returns, with *g NULL on failure.
 * mr3_graph_launch replays it on stream (NULL: the library's per-thread stream),
 * waits, and fills the problems' scalar outputs like mr3_solve_batch; it returns 4,
 * with info 2 for the pipeline's problems, when their tree this time needs more
 * levels or slots than the capture had (new D and E): solve them with
 * mr3_solve_batch, or capture again.  probs must be the array given at creation.
 */
typedef struct mr3_graph mr3_graph;
int mr3_graph_create(mr3_problem *probs, int nprob, void *stream, mr3_graph **g);
int mr3_graph_launch(mr3_graph *g, mr3_problem *probs, void *stream);
void mr3_graph_destroy(mr3_graph *g);

/*
 * Serving: a request loop that admits problems as device memory and concurrency
 * slots free up, runs them on several streams at once and reports completion.
 * budget_bytes bounds the device memory of admitted problems (8 n m for Z plus the
 * per-problem workspace); max_inflight bounds the number of batches on the device.
 * Problems use host arrays; the server copies them in and results out.
 */
typedef struct mr3_server mr3_server;

typedef struct mr3_request {
    int n;
    char jobz, range;
    char pad0, pad1;
    double vl, vu;
    int il, iu, tryrac;
    const double *D, *E;   /* host inputs, must stay valid until completion   */
    double *W;             /* host output, n entries                           */
    double *Z;             /* host output, n x (columns) with ldz = n, or NULL */
    int *Zsupp;            /* host output, 2n ints, or NULL                    */
    int mcols;             /* columns available in Z                           */
    /* outputs */
    int m, info;
    double t_submit, t_start, t_done;   /* seconds, steady clock */
} mr3_request;

mr3_server *mr3_server_create(long long budget_bytes, int max_inflight, int max_batch);
/* Queues a request; returns its id (>= 0) or -1 if it can never fit the budget. */
long long mr3_server_submit(mr3_server *s, mr3_request *req);
/* Blocks until request id is complete; returns its info (0, or 1 invalid argument, 2
 * failure), also when it completed before the call, and -1 for an id that is unknown or
 * was already waited for.  A failure of the group's device memory or copies fails every
 * request of the group; a solver failure only its own request.  A range-V request
 * without mcols reserves its interval's eigenvalue count plus 4 columns on the device
 * (the count the reference's jobz "C" query returns) and is solved again with n
 * columns if the interval holds more. */
int mr3_server_wait(mr3_server *s, long long id);
/* Blocks until every submitted request is complete. */
void mr3_server_drain(mr3_server *s);
void mr3_server_destroy(mr3_server *s);
/* Monotonic clock used for the request timestamps. */
double mr3_now(void);

/*
 * Profiling.  With on = 1, every kernel launch of the library is bracketed by HIP
 * events on the stream it runs on and the kernels count their algorithmic work
 * (sweep elements per kernel family); on = 2 records the events only (the kernels
 * then run exactly as with profiling off); on = 0 turns both off.
 * mr3_profile_report writes a JSON object
 * {"kernels": {name: {"ms", "launches"}}, "steps": {family: elements}} into buf and
 * returns its length (or minus the needed size if buf is too small).  No counterpart
 * in the reference, which has no instrumentation.
 */
void mr3_profile_enable(int on);
void mr3_profile_reset(void);
int mr3_profile_report(char *buf, int len);

/*
 * Launch-shape parameters, process-wide, read at the start of every solve.  They
 * change how the work is spread over the device, never the results.  No counterpart
 * in the reference (whose only knob is PMR_NUM_THREADS); the defaults were measured on
 * the batched workload, and mr3smp_amd/tune.py searches them for another workload.
 *   side_streams     streams for the singleton phases, 1..3 (0: 3 when
 *                    GPU_MAX_HW_QUEUES >= 8, else 2)
 *   tail_singletons  iterations with at most this many singletons going on run in the
 *                    latency kernel (one workgroup per singleton), 0..2^24
 *   group_lanes      lane budget of the multisection cluster refinements, >= 4096
 *   rrf_group_max    levels with more clusters run dlarrf on one lane per cluster
 *   singleton_waves  most waves of one singleton launch, >= 64
 *   small_first      matrices of the small route (n < 500) of at least this order start
 *                    with the solve, the others once the large problems' root
 *                    eigenvalues are refined; 0: all at once.  0..2^20
 * mr3_set_tuning returns 0, or 1 (and changes nothing) when a value is out of range.
 */
typedef struct mr3_tuning {
    int side_streams;
    int tail_singletons;
    int group_lanes;
    int rrf_group_max;
    int singleton_waves;
    int small_first;
} mr3_tuning;
void mr3_get_tuning(mr3_tuning *t);
int mr3_set_tuning(const mr3_tuning *t);

/* Library version string. */
const char *mr3_version(void);

#ifdef __cplusplus
}
#endif
#endif
