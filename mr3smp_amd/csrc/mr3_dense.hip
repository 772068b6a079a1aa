// mr3_dense.hip -- dense symmetric eigensolvers on top of the tridiagonal MRRR solver
// (HPAC/mr3smp SRC/dense.c:151-316, INCLUDE/mrrr.h:270-372).
//
// dsyeig: A is scaled into the safe range (dense.c:476-526), reduced to tridiagonal
// form by rocSOLVER's dsytrd, solved by the batched MRRR pipeline on the device, and
// the eigenvectors are back-transformed by rocSOLVER's dormtr.  A, the tridiagonal and
// Z stay in HBM from the reduction to the back-transformation.  dsygeig adds the
// Cholesky factorisation and the reduction to standard form (rocSOLVER potrf, sygst)
// and the back-transformation of the vectors (rocBLAS trsm or trmm), as dense.c:236-316.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "../../include/mr3hip.h"

namespace {

// Largest |A(i,j)| over the stored triangle (dlansy 'M'); |x| >= 0 orders like its
// bit pattern read as an unsigned integer.
__global__ __launch_bounds__(256) void k_dense_maxabs(const double* __restrict__ A, int lda, int n,
                                                      int lower, unsigned long long* out) {
    unsigned long long best = 0;
    const long long total = (long long)n * n;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        if (lower ? i >= j : i <= j) {
            const double v = fabs(A[(size_t)j * lda + i]);
            const unsigned long long b = (unsigned long long)__double_as_longlong(v);
            best = b > best ? b : best;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(best, o);
        best = t > best ? t : best;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

// The stored triangle times sigma (the dscal loops of dense.c:505-516).
__global__ __launch_bounds__(256) void k_dense_scale(double* A, int lda, int n, int lower,
                                                     double sigma) {
    const long long total = (long long)n * n;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        if (lower ? i >= j : i <= j) A[(size_t)j * lda + i] *= sigma;
    }
}

__global__ void k_set(double* p, double v) { *p = v; }

struct Ctx {
    hipStream_t st = nullptr;
    rocblas_handle h = nullptr;
    bool ok = false;
};

// One stream and rocBLAS handle per host thread, kept for the thread's lifetime.
Ctx& ctx() {
    thread_local Ctx c;
    if (!c.ok && !c.h) {
        if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) == hipSuccess &&
            rocblas_create_handle(&c.h) == rocblas_status_success &&
            rocblas_set_stream(c.h, c.st) == rocblas_status_success)
            c.ok = true;
    }
    return c;
}

inline int up(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }

// Device buffers of one call, released on the call's stream.
struct Buf {
    hipStream_t st;
    void* p[8] = {};
    int k = 0;
    explicit Buf(hipStream_t s) : st(s) {}
    template <class T>
    T* get(size_t count) {
        void* q = nullptr;
        if (k == 8 || hipMallocAsync(&q, (count ? count : 1) * sizeof(T), st) != hipSuccess) return nullptr;
        p[k++] = q;
        return (T*)q;
    }
    ~Buf() {
        (void)hipStreamSynchronize(st);
        for (int i = 0; i < k; ++i) (void)hipFreeAsync(p[i], st);
        (void)hipStreamSynchronize(st);
    }
};

// Argument checks of dsyeig (dense.c:174-183), plus the storage arguments.
int check_syeig(int jz, int rg, int ul, int n, int lda, double vl, double vu, int il, int iu) {
    if (!(jz == 'N' || jz == 'V' || jz == 'C')) return 1;
    if (!(rg == 'A' || rg == 'V' || rg == 'I')) return 1;
    if (!(ul == 'L' || ul == 'U')) return 1;
    if (n <= 0 || lda < n) return 1;
    if (rg == 'V') {
        if (vu <= vl) return 1;
    } else if (rg == 'I') {
        if (il < 1 || il > n || iu < il || iu > n) return 1;
    }
    return 0;
}

}  // namespace

// The dense symmetric problem on device memory: dA (lda x n, the uplo triangle) is
// overwritten by the reduction; W gets m eigenvalues, Z (ldz x m) the eigenvectors.
extern "C" int mr3_dsyeig_device(char jobz, char range, char uplo, int n, double* dA, int lda,
                                 double* vl, double* vu, int* il, int* iu, int* m, double* dW,
                                 double* dZ, int ldz, int zcols, void* stream) {
    const int jz = up(jobz), rg = up(range), ul = up(uplo);
    *m = 0;
    if (check_syeig(jz, rg, ul, n, lda, rg == 'V' ? *vl : 0.0, rg == 'V' ? *vu : 1.0,
                    rg == 'I' ? *il : 1, rg == 'I' ? *iu : 1))
        return 1;
    if (jz == 'V' && (dZ == nullptr || ldz < n)) return 1;
    if (jz == 'C' && rg != 'V') {   // the count needs no reduction (mrrr.c:130-138)
        *m = (rg == 'A') ? n : *iu - *il + 1;
        return 0;
    }
    Ctx& c = ctx();
    if (!c.ok) return 2;
    hipStream_t st = stream ? (hipStream_t)stream : c.st;
    if (rocblas_set_stream(c.h, st) != rocblas_status_success) return 2;
    const bool lower = ul == 'L';
    const rocblas_fill fill = lower ? rocblas_fill_lower : rocblas_fill_upper;
    Buf B(st);
    double* dD = B.get<double>(n);
    double* dE = B.get<double>(n);
    double* dTau = B.get<double>(n);
    unsigned long long* dmax = B.get<unsigned long long>(1);
    int* dZs = B.get<int>(2 * (size_t)n);
    if (!dD || !dE || !dTau || !dmax || !dZs) return 2;
    // scale into [rmin, rmax] (dense.c:476-526)
    const double smlnum = DBL_MIN / DBL_EPSILON, bignum = 1.0 / smlnum;
    const double rmin = sqrt(smlnum), rmax = fmin(sqrt(bignum), 1.0 / sqrt(sqrt(DBL_MIN)));
    unsigned long long hmax = 0;
    const unsigned grid = (unsigned)std::min<long long>(((long long)n * n + 255) / 256, 4096);
    if (hipMemsetAsync(dmax, 0, sizeof(unsigned long long), st) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_dense_maxabs, dim3(grid), dim3(256), 0, st, dA, lda, n, lower ? 1 : 0, dmax);
    if (hipMemcpyAsync(&hmax, dmax, sizeof(hmax), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return 2;
    double norm;
    memcpy(&norm, &hmax, sizeof(norm));
    double sigma = 1.0;
    bool scaled = false;
    if (norm > 0.0 && norm < rmin) { scaled = true; sigma = rmin / norm; }
    else if (norm > rmax) { scaled = true; sigma = rmax / norm; }
    double svl = *vl, svu = *vu;
    if (scaled) {
        hipLaunchKernelGGL(k_dense_scale, dim3(grid), dim3(256), 0, st, dA, lda, n, lower ? 1 : 0, sigma);
        if (rg == 'V') { svl *= sigma; svu *= sigma; }
    }
    // reduction to tridiagonal form; E(n) is not part of T
    if (rocsolver_dsytrd(c.h, fill, n, dA, lda, dD, dE, dTau) != rocblas_status_success) return 2;
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, st, dE + (n - 1), 0.0);
    // the tridiagonal problem (mrrr, dense.c:203-205)
    mr3_problem p;
    memset(&p, 0, sizeof(p));
    p.n = n; p.jobz = (char)jz; p.range = (char)rg;
    p.vl = svl; p.vu = svu;
    p.il = *il;
    p.iu = *iu;
    p.tryrac = 1;
    p.D = dD; p.E = dE; p.W = dW;
    p.Z = (jz == 'V') ? dZ : nullptr;
    p.ldz = ldz;
    p.mcols = (jz == 'V') ? zcols : 0;
    p.Zsupp = (jz == 'V') ? dZs : nullptr;
    int rc = mr3_solve_batch(&p, 1, st);
    if (rc != 0 || p.info != 0) return rc ? rc : p.info;
    *m = p.m;
    // back-transformation Z = Q Z (dense.c:208-211)
    if (jz == 'V' && p.m > 0) {
        if (rocsolver_dormtr(c.h, rocblas_side_left, fill, rocblas_operation_none, n, p.m, dA, lda,
                             dTau, dZ, ldz) != rocblas_status_success)
            return 2;
    }
    // unscale the eigenvalues and the interval (dense.c:214-220)
    const double inv = 1.0 / sigma;
    if (scaled && p.m > 0 && jz != 'C') {
        if (rocblas_dscal(c.h, p.m, &inv, dW, 1) != rocblas_status_success) return 2;
    }
    *vl = p.vl_out; *vu = p.vu_out; *il = p.il_out; *iu = p.iu_out;
    if (scaled) { *vl *= inv; *vu *= inv; }
    if (hipStreamSynchronize(st) != hipSuccess) return 2;
    return 0;
}

// dsyeig (INCLUDE/mrrr.h:270-272, SRC/dense.c:151-230): host arrays in and out.
extern "C" int dsyeig(char* jobz, char* range, char* uplo, int* np, double* A, int* ldap,
                      double* vl, double* vu, int* il, int* iu, int* m, double* W, double* Z,
                      int* ldz) {
    const int n = *np, lda = *ldap;
    const int jz = up(jobz[0]), rg = up(range[0]), ul = up(uplo[0]);
    if (check_syeig(jz, rg, ul, n, lda, rg == 'V' ? *vl : 0.0, rg == 'V' ? *vu : 1.0,
                    rg == 'I' ? *il : 1, rg == 'I' ? *iu : 1))
        return 1;
    if (jz == 'V' && *ldz < n) return 1;
    Ctx& c = ctx();
    if (!c.ok) return 2;
    hipStream_t st = c.st;
    const int zcols = (rg == 'I') ? *iu - *il + 1 : n;
    Buf B(st);
    double* dA = B.get<double>((size_t)lda * n);
    double* dW = B.get<double>(n);
    double* dZ = (jz == 'V') ? B.get<double>((size_t)n * zcols) : nullptr;
    if (!dA || !dW || (jz == 'V' && !dZ)) return 2;
    if (hipMemcpyAsync(dA, A, sizeof(double) * (size_t)lda * n, hipMemcpyHostToDevice, st) != hipSuccess)
        return 2;
    int rc = mr3_dsyeig_device((char)jz, (char)rg, (char)ul, n, dA, lda, vl, vu, il, iu, m, dW, dZ,
                               n, zcols, st);
    if (rc) return rc;
    if (*m > 0 && jz != 'C') {
        if (hipMemcpyAsync(W, dW, sizeof(double) * *m, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
        if (jz == 'V' &&
            hipMemcpy2DAsync(Z, sizeof(double) * (size_t)*ldz, dZ, sizeof(double) * (size_t)n,
                             sizeof(double) * (size_t)n, *m, hipMemcpyDeviceToHost, st) != hipSuccess)
            return 2;
    }
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

// dsygeig (INCLUDE/mrrr.h:352-356, SRC/dense.c:236-316): A x = lambda B x (itype 1),
// A B x = lambda x (2) or B A x = lambda x (3); eigenvectors returned in A, B
// overwritten by its Cholesky factor.
extern "C" int dsygeig(int* itype, char* jobz, char* range, char* uplo, int* np, double* A,
                       int* ldap, double* B, int* ldbp, double* vl, double* vu, int* il, int* iu,
                       int* m, double* W) {
    const int n = *np, lda = *ldap, ldb = *ldbp, it = *itype;
    const int jz = up(jobz[0]), rg = up(range[0]), ul = up(uplo[0]);
    if (n <= 1 || it < 1 || it > 3) return 1;   // dense.c:254-256
    if (check_syeig(jz, rg, ul, n, lda, rg == 'V' ? *vl : 0.0, rg == 'V' ? *vu : 1.0,
                    rg == 'I' ? *il : 1, rg == 'I' ? *iu : 1) || ldb < n)
        return 1;
    Ctx& c = ctx();
    if (!c.ok) return 2;
    hipStream_t st = c.st;
    if (rocblas_set_stream(c.h, st) != rocblas_status_success) return 2;
    const rocblas_fill fill = (ul == 'L') ? rocblas_fill_lower : rocblas_fill_upper;
    const int zcols = (rg == 'I') ? *iu - *il + 1 : n;
    Buf Bf(st);
    double* dA = Bf.get<double>((size_t)lda * n);
    double* dB = Bf.get<double>((size_t)ldb * n);
    double* dW = Bf.get<double>(n);
    double* dZ = Bf.get<double>((size_t)n * zcols);
    double* dT = (jz == 'V' && it == 3) ? Bf.get<double>((size_t)n * zcols) : nullptr;
    rocblas_int* dinfo = Bf.get<rocblas_int>(1);
    if (!dA || !dB || !dW || !dZ || !dinfo || (jz == 'V' && it == 3 && !dT)) return 2;
    if (hipMemcpyAsync(dA, A, sizeof(double) * (size_t)lda * n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dB, B, sizeof(double) * (size_t)ldb * n, hipMemcpyHostToDevice, st) != hipSuccess)
        return 2;
    // Cholesky factor of B (dense.c:265-267)
    rocblas_int hinfo = 0;
    if (rocsolver_dpotrf(c.h, fill, n, dB, ldb, dinfo) != rocblas_status_success ||
        hipMemcpyAsync(&hinfo, dinfo, sizeof(hinfo), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return 2;
    if (hinfo != 0) return 2;   // B is not positive definite
    // standard form (dense.c:269-271)
    const rocblas_eform ef = it == 1 ? rocblas_eform_ax : (it == 2 ? rocblas_eform_abx : rocblas_eform_bax);
    if (rocsolver_dsygst(c.h, ef, fill, n, dA, lda, dB, ldb) != rocblas_status_success) return 2;
    int rc = mr3_dsyeig_device((char)jz, (char)rg, (char)ul, n, dA, lda, vl, vu, il, iu, m, dW, dZ,
                               n, zcols, st);
    if (rc) return rc;
    // back-transformation (dense.c:278-300)
    const double one = 1.0;
    double* dX = dZ;
    if (jz == 'V' && *m > 0) {
        if (it == 1 || it == 2) {
            const rocblas_operation tr = (ul == 'U') ? rocblas_operation_none : rocblas_operation_transpose;
            if (rocblas_dtrsm(c.h, rocblas_side_left, fill, tr, rocblas_diagonal_non_unit, n, *m, &one,
                              dB, ldb, dZ, n) != rocblas_status_success)
                return 2;
        } else {
            const rocblas_operation tr = (ul == 'U') ? rocblas_operation_transpose : rocblas_operation_none;
            if (rocblas_dtrmm(c.h, rocblas_side_left, fill, tr, rocblas_diagonal_non_unit, n, *m, &one,
                              dB, ldb, dZ, n, dT, n) != rocblas_status_success)
                return 2;
            dX = dT;
        }
    }
    if (*m > 0 && jz != 'C') {
        if (hipMemcpyAsync(W, dW, sizeof(double) * *m, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
        // eigenvectors into the first m columns of A (dense.c:303-305)
        if (jz == 'V' &&
            hipMemcpy2DAsync(A, sizeof(double) * (size_t)lda, dX, sizeof(double) * (size_t)n,
                             sizeof(double) * (size_t)n, *m, hipMemcpyDeviceToHost, st) != hipSuccess)
            return 2;
    }
    // B holds its Cholesky factor on return, as in the reference
    if (hipMemcpyAsync(B, dB, sizeof(double) * (size_t)ldb * n, hipMemcpyDeviceToHost, st) != hipSuccess)
        return 2;
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

// Fortran bindings (SRC/dense.c:1037-1054).
extern "C" void dsyeig_(char* jobz, char* range, char* uplo, int* n, double* A, int* lda,
                        double* vl, double* vu, int* il, int* iu, int* m, double* W, double* Z,
                        int* ldz, int* info) {
    *info = dsyeig(jobz, range, uplo, n, A, lda, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void dsygeig_(int* itype, char* jobz, char* range, char* uplo, int* np, double* A,
                         int* ldap, double* B, int* ldbp, double* vlp, double* vup, int* ilp,
                         int* iup, int* mp, double* W, int* info) {
    *info = dsygeig(itype, jobz, range, uplo, np, A, ldap, B, ldbp, vlp, vup, ilp, iup, mp, W);
}
