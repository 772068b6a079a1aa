// mr3_dense.hip -- dense symmetric / Hermitian eigensolvers on top of the tridiagonal
// MRRR solver (HPAC/mr3smp SRC/dense.c:151-1114, INCLUDE/mrrr.h:270-806).
//
// The standard problem: the matrix is scaled into the safe range (dense.c:476-566,
// 940-1034), reduced to real tridiagonal form by rocSOLVER (sytrd / hetrd), solved by
// the batched MRRR pipeline on the device, and the eigenvectors are back-transformed by
// rocSOLVER (ormtr / unmtr).  The generalized problem adds the Cholesky factorisation
// and the reduction to standard form (potrf, sygst / hegst) and the back-transformation
// of the vectors (rocBLAS trsm or trmm).  Matrix, tridiagonal and eigenvectors stay in
// HBM from the reduction to the back-transformation.
//
// Packed storage (dspeig, dspgeig, zhpeig, zhpgeig) is unpacked on the device into a
// full-storage work matrix, since the reduction and back-transformation are the same
// mathematics on either storage; the Cholesky factor a packed generalized call returns
// is packed again on the device.
//
// The reference overwrites A (or AP) with the Householder vectors of the reduction and
// frees their scalar factors TAU, so a caller cannot use what A holds on return; these
// routines leave A unmodified in the standard problem.  In the generalized problem A
// receives the eigenvectors and B its Cholesky factor, as in the reference.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "../../include/mr3hip.h"
#include "mr3_staging.h"

namespace {

using zc = rocblas_double_complex;

__device__ __forceinline__ double mag(double x) { return fabs(x); }
__device__ __forceinline__ double mag(zc x) { return hypot(x.real(), x.imag()); }
__device__ __forceinline__ double mag_diag(double x) { return fabs(x); }
__device__ __forceinline__ double mag_diag(zc x) { return fabs(x.real()); }   // zlanhe 'M'
__device__ __forceinline__ double scaled(double x, double s) { return x * s; }
__device__ __forceinline__ zc scaled(zc x, double s) { return zc(x.real() * s, x.imag() * s); }

// Column-major position of A(i, j) in packed storage (i <= j for upper, i >= j for lower).
__device__ __forceinline__ long long packed_at(int i, int j, int n, int lower) {
    return lower ? (long long)i + (long long)j * (2LL * n - j - 1) / 2
                 : (long long)i + (long long)j * (j + 1) / 2;
}

// Largest magnitude over the stored triangle (dlansy / zlanhe 'M'); a magnitude >= 0
// orders like its bit pattern read as an unsigned integer.
template <class T>
__global__ __launch_bounds__(256) void k_dense_maxabs(const T* __restrict__ A, int lda, int n,
                                                      int lower, unsigned long long* out) {
    unsigned long long best = 0;
    const long long total = (long long)n * n;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        if (lower ? i >= j : i <= j) {
            const T a = A[(size_t)j * lda + i];
            const double v = (i == j) ? mag_diag(a) : mag(a);
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

// The stored triangle times sigma (dense.c:505-516, 960-975).
template <class T>
__global__ __launch_bounds__(256) void k_dense_scale(T* A, int lda, int n, int lower, double sigma) {
    const long long total = (long long)n * n;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        if (lower ? i >= j : i <= j) A[(size_t)j * lda + i] = scaled(A[(size_t)j * lda + i], sigma);
    }
}

// Packed triangle -> the same triangle of a full n x n matrix (leading dimension n), or
// back (to_packed = 1).
template <class T>
__global__ __launch_bounds__(256) void k_dense_packed(T* AP, T* A, int n, int lower, int to_packed) {
    const long long total = (long long)n * n;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        if (lower ? i >= j : i <= j) {
            const long long p = packed_at(i, j, n, lower);
            if (to_packed) AP[p] = A[(size_t)j * n + i];
            else A[(size_t)j * n + i] = AP[p];
        }
    }
}

// Real eigenvectors (n x m, leading dimension n) -> complex columns (leading dimension ldz).
__global__ __launch_bounds__(256) void k_dense_to_complex(const double* __restrict__ Zr, zc* Z,
                                                          int n, int m, int ldz) {
    const long long total = (long long)n * m;
    for (long long k = blockIdx.x * 256LL + threadIdx.x; k < total; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / n), i = (int)(k % n);
        Z[(size_t)j * ldz + i] = zc(Zr[(size_t)j * n + i], 0.0);
    }
}

__global__ void k_set(double* p, double v) { *p = v; }

inline unsigned grid_for(long long elems) {
    return (unsigned)std::max<long long>(1, std::min<long long>((elems + 255) / 256, 4096));
}

// The library routines of each arithmetic.
template <class T> struct Lin;
template <> struct Lin<double> {
    static constexpr rocblas_operation adj = rocblas_operation_transpose;
    static rocblas_status trd(rocblas_handle h, rocblas_fill f, int n, double* A, int lda, double* D,
                              double* E, double* tau) {
        return rocsolver_dsytrd(h, f, n, A, lda, D, E, tau);
    }
    static rocblas_status mtr(rocblas_handle h, rocblas_fill f, int n, int m, double* A, int lda,
                              double* tau, double* C, int ldc) {
        return rocsolver_dormtr(h, rocblas_side_left, f, rocblas_operation_none, n, m, A, lda, tau, C, ldc);
    }
    static rocblas_status potrf(rocblas_handle h, rocblas_fill f, int n, double* B, int ldb,
                                rocblas_int* info) {
        return rocsolver_dpotrf(h, f, n, B, ldb, info);
    }
    static rocblas_status gst(rocblas_handle h, rocblas_eform e, rocblas_fill f, int n, double* A,
                              int lda, double* B, int ldb) {
        return rocsolver_dsygst(h, e, f, n, A, lda, B, ldb);
    }
    static rocblas_status trsm(rocblas_handle h, rocblas_fill f, rocblas_operation t, int n, int m,
                               const double* B, int ldb, double* X, int ldx) {
        const double one = 1.0;
        return rocblas_dtrsm(h, rocblas_side_left, f, t, rocblas_diagonal_non_unit, n, m, &one, B, ldb,
                             X, ldx);
    }
    static rocblas_status trmm(rocblas_handle h, rocblas_fill f, rocblas_operation t, int n, int m,
                               const double* B, int ldb, const double* X, int ldx, double* Y, int ldy) {
        const double one = 1.0;
        return rocblas_dtrmm(h, rocblas_side_left, f, t, rocblas_diagonal_non_unit, n, m, &one, B, ldb,
                             X, ldx, Y, ldy);
    }
};
template <> struct Lin<zc> {
    static constexpr rocblas_operation adj = rocblas_operation_conjugate_transpose;
    static rocblas_status trd(rocblas_handle h, rocblas_fill f, int n, zc* A, int lda, double* D,
                              double* E, zc* tau) {
        return rocsolver_zhetrd(h, f, n, A, lda, D, E, tau);
    }
    static rocblas_status mtr(rocblas_handle h, rocblas_fill f, int n, int m, zc* A, int lda, zc* tau,
                              zc* C, int ldc) {
        return rocsolver_zunmtr(h, rocblas_side_left, f, rocblas_operation_none, n, m, A, lda, tau, C, ldc);
    }
    static rocblas_status potrf(rocblas_handle h, rocblas_fill f, int n, zc* B, int ldb,
                                rocblas_int* info) {
        return rocsolver_zpotrf(h, f, n, B, ldb, info);
    }
    static rocblas_status gst(rocblas_handle h, rocblas_eform e, rocblas_fill f, int n, zc* A, int lda,
                              zc* B, int ldb) {
        return rocsolver_zhegst(h, e, f, n, A, lda, B, ldb);
    }
    static rocblas_status trsm(rocblas_handle h, rocblas_fill f, rocblas_operation t, int n, int m,
                               const zc* B, int ldb, zc* X, int ldx) {
        const zc one(1.0, 0.0);
        return rocblas_ztrsm(h, rocblas_side_left, f, t, rocblas_diagonal_non_unit, n, m, &one, B, ldb,
                             X, ldx);
    }
    static rocblas_status trmm(rocblas_handle h, rocblas_fill f, rocblas_operation t, int n, int m,
                               const zc* B, int ldb, const zc* X, int ldx, zc* Y, int ldy) {
        const zc one(1.0, 0.0);
        return rocblas_ztrmm(h, rocblas_side_left, f, t, rocblas_diagonal_non_unit, n, m, &one, B, ldb,
                             X, ldx, Y, ldy);
    }
};

struct Ctx {
    hipStream_t st = nullptr;
    rocblas_handle h = nullptr;
    bool ok = false;
};

// One stream and rocBLAS handle per host thread, kept for the thread's lifetime.  Atomic
// accumulation is off so that a call gives the same bits every time, as the reference does.
Ctx& ctx() {
    thread_local Ctx c;
    thread_local bool tried = false;
    if (!tried) {
        tried = true;
        if (hipStreamCreateWithFlags(&c.st, hipStreamNonBlocking) == hipSuccess &&
            rocblas_create_handle(&c.h) == rocblas_status_success &&
            rocblas_set_stream(c.h, c.st) == rocblas_status_success &&
            rocblas_set_atomics_mode(c.h, rocblas_atomics_not_allowed) == rocblas_status_success)
            c.ok = true;
    }
    return c;
}

inline int up(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }

// Device buffers of one call, released on the call's stream when it returns.
struct Buf {
    hipStream_t st;
    void* p[10] = {};
    int k = 0;
    explicit Buf(hipStream_t s) : st(s) {}
    template <class T>
    T* get(size_t count) {
        void* q = nullptr;
        if (k == 10 || hipMallocAsync(&q, (count ? count : 1) * sizeof(T), st) != hipSuccess) return nullptr;
        p[k++] = q;
        return (T*)q;
    }
    ~Buf() {
        (void)hipStreamSynchronize(st);
        for (int i = 0; i < k; ++i) (void)hipFreeAsync(p[i], st);
        (void)hipStreamSynchronize(st);
    }
};

// Columns 0 .. m-1 of a device block (leading dimension n) into host memory with
// leading dimension ld, large blocks through the pinned staging buffers.
template <class T>
int copy_block_out(T* dst, int ld, const T* src, int n, int m, hipStream_t st) {
    constexpr size_t w = sizeof(T) / sizeof(double);
    const bool big = sizeof(T) * (size_t)n * m > mr3::kStageDirect;
    mr3::Staging* sg = big ? mr3::staging_get() : nullptr;
    mr3::Staging direct;
    const int rc = mr3::copy_cols_out((double*)dst, w * (size_t)ld, (const double*)src, w * (size_t)n,
                                      (size_t)m, st, sg ? *sg : direct);
    if (sg) mr3::staging_put(sg);
    return rc;
}

// Argument checks of the standard drivers (dense.c:174-180, 341-347, 596-602, 788-794).
int check_std(int jz, int rg, int ul, int n, const double* vl, const double* vu, const int* il,
              const int* iu) {
    if (!(jz == 'N' || jz == 'V' || jz == 'C')) return 1;
    if (!(rg == 'A' || rg == 'V' || rg == 'I')) return 1;
    if (!(ul == 'L' || ul == 'U')) return 1;
    if (n <= 0) return 1;
    if (rg == 'V') {
        if (*vu <= *vl) return 1;
    } else if (rg == 'I') {
        if (*il < 1 || *il > n || *iu < *il || *iu > n) return 1;
    }
    return 0;
}

// Columns the eigenvectors may need (the reference sizes Z alike, dense.c:262-263).
int zcols_for(int rg, int n, const int* il, const int* iu) { return rg == 'I' ? *iu - *il + 1 : n; }

// The standard problem on a device matrix dA (lda x n, uplo triangle), overwritten by the
// reduction.  dW receives m eigenvalues, dZ (ldz x zcols) the eigenvectors.
template <class T>
int standard_device(int jz, int rg, int ul, int n, T* dA, int lda, double* vl, double* vu, int* il,
                    int* iu, int* m, double* dW, T* dZ, int ldz, int zcols, hipStream_t st, Ctx& c) {
    *m = 0;
    if (jz == 'C' && rg != 'V') {   // the count needs no reduction (mrrr.c:130-138)
        *m = (rg == 'A') ? n : *iu - *il + 1;
        return 0;
    }
    if (rocblas_set_stream(c.h, st) != rocblas_status_success) return 2;
    const bool lower = ul == 'L';
    const rocblas_fill fill = lower ? rocblas_fill_lower : rocblas_fill_upper;
    constexpr bool cplx = !std::is_same<T, double>::value;
    Buf B(st);
    double* dD = B.get<double>(n);
    double* dE = B.get<double>(n);
    T* dTau = B.get<T>(n);
    unsigned long long* dmax = B.get<unsigned long long>(1);
    int* dZs = B.get<int>(2 * (size_t)n);
    double* dZr = (cplx && jz == 'V') ? B.get<double>((size_t)n * zcols) : nullptr;
    if (!dD || !dE || !dTau || !dmax || !dZs || (cplx && jz == 'V' && !dZr)) return 2;
    // scale into [rmin, rmax] (dense.c:476-526)
    const double smlnum = DBL_MIN / DBL_EPSILON, bignum = 1.0 / smlnum;
    const double rmin = sqrt(smlnum), rmax = fmin(sqrt(bignum), 1.0 / sqrt(sqrt(DBL_MIN)));
    const unsigned grid = grid_for((long long)n * n);
    unsigned long long hmax = 0;
    if (hipMemsetAsync(dmax, 0, sizeof(unsigned long long), st) != hipSuccess) return 2;
    hipLaunchKernelGGL(k_dense_maxabs<T>, dim3(grid), dim3(256), 0, st, dA, lda, n, lower ? 1 : 0, dmax);
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
        hipLaunchKernelGGL(k_dense_scale<T>, dim3(grid), dim3(256), 0, st, dA, lda, n, lower ? 1 : 0, sigma);
        if (rg == 'V') { svl *= sigma; svu *= sigma; }
    }
    // reduction to real tridiagonal form; E(n) is not part of T
    if (Lin<T>::trd(c.h, fill, n, dA, lda, dD, dE, dTau) != rocblas_status_success) return 2;
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, st, dE + (n - 1), 0.0);
    // the tridiagonal problem (dense.c:205-207)
    mr3_problem p;
    memset(&p, 0, sizeof(p));
    p.n = n; p.jobz = (char)jz; p.range = (char)rg;
    p.vl = svl; p.vu = svu; p.il = *il; p.iu = *iu;
    p.tryrac = 1;
    p.D = dD; p.E = dE; p.W = dW;
    if (jz == 'V') {
        p.Z = cplx ? dZr : (double*)dZ;
        p.ldz = cplx ? n : ldz;
        p.mcols = zcols;
        p.Zsupp = dZs;
    }
    const int rc = mr3_solve_batch(&p, 1, st);
    if (rc != 0 || p.info != 0) return rc ? rc : p.info;
    *m = p.m;
    // back-transformation Z = Q Z (dense.c:211-213, 646-653)
    if (jz == 'V' && p.m > 0) {
        if constexpr (cplx) {
            hipLaunchKernelGGL(k_dense_to_complex, dim3(grid_for((long long)n * p.m)), dim3(256), 0, st,
                               dZr, (zc*)dZ, n, p.m, ldz);
        }
        if (Lin<T>::mtr(c.h, fill, n, p.m, dA, lda, dTau, dZ, ldz) != rocblas_status_success) return 2;
    }
    // unscale the eigenvalues and the interval (dense.c:216-222)
    const double inv = 1.0 / sigma;
    if (scaled && p.m > 0 && jz != 'C') {
        if (rocblas_dscal(c.h, p.m, &inv, dW, 1) != rocblas_status_success) return 2;
    }
    *vl = p.vl_out; *vu = p.vu_out; *il = p.il_out; *iu = p.iu_out;
    if (scaled) { *vl *= inv; *vu *= inv; }
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

// The standard problem on host arrays: A full (lda > 0) or packed (lda = 0); results
// into W and Z (ldz).
template <class T>
int standard_host(int jz, int rg, int ul, int n, const T* A, int lda, double* vl, double* vu, int* il,
                  int* iu, int* m, double* W, T* Z, int ldz) {
    *m = 0;
    if (check_std(jz, rg, ul, n, vl, vu, il, iu)) return 1;
    if (lda != 0 && lda < n) return 1;
    if (jz == 'V' && (Z == nullptr || ldz < n)) return 1;
    if (jz == 'C' && rg != 'V') { *m = (rg == 'A') ? n : *iu - *il + 1; return 0; }
    Ctx& c = ctx();
    if (!c.ok) return 2;
    hipStream_t st = c.st;
    const int zcols = zcols_for(rg, n, il, iu);
    Buf B(st);
    T* dA = B.get<T>((size_t)n * n);
    double* dW = B.get<double>(n);
    T* dZ = (jz == 'V') ? B.get<T>((size_t)n * zcols) : nullptr;
    T* dAP = (lda == 0) ? B.get<T>((size_t)n * (n + 1) / 2) : nullptr;
    if (!dA || !dW || (jz == 'V' && !dZ) || (lda == 0 && !dAP)) return 2;
    if (lda == 0) {
        if (hipMemcpyAsync(dAP, A, sizeof(T) * (size_t)n * (n + 1) / 2, hipMemcpyHostToDevice, st) !=
            hipSuccess)
            return 2;
        hipLaunchKernelGGL(k_dense_packed<T>, dim3(grid_for((long long)n * n)), dim3(256), 0, st, dAP, dA, n,
                           ul == 'L' ? 1 : 0, 0);
    } else if (hipMemcpy2DAsync(dA, sizeof(T) * (size_t)n, A, sizeof(T) * (size_t)lda, sizeof(T) * (size_t)n,
                                n, hipMemcpyHostToDevice, st) != hipSuccess) {
        return 2;
    }
    const int rc = standard_device<T>(jz, rg, ul, n, dA, n, vl, vu, il, iu, m, dW, dZ, n, zcols, st, c);
    if (rc) return rc;
    if (*m > 0 && jz != 'C') {
        if (hipMemcpyAsync(W, dW, sizeof(double) * *m, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
        if (jz == 'V' && copy_block_out(Z, ldz, dZ, n, *m, st) != 0) return 2;
    }
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

// The generalized problem on host arrays (dense.c:236-316, 403-471, 676-757, 869-935).
// Full storage: lda, ldb > 0, eigenvectors into A's first m columns.  Packed storage:
// lda = ldb = 0, eigenvectors into Z (ldz).  B gets its Cholesky factor either way.
template <class T>
int general_host(int it, int jz, int rg, int ul, int n, T* A, int lda, T* Bm, int ldb, double* vl,
                 double* vu, int* il, int* iu, int* m, double* W, T* Z, int ldz) {
    *m = 0;
    const bool packed = (lda == 0);
    if (it < 1 || it > 3) return 1;
    if (check_std(jz, rg, ul, n, vl, vu, il, iu)) return 1;
    if (!packed && (lda < n || ldb < n)) return 1;
    if (packed && jz == 'V' && (Z == nullptr || ldz < n)) return 1;
    Ctx& c = ctx();
    if (!c.ok) return 2;
    hipStream_t st = c.st;
    if (rocblas_set_stream(c.h, st) != rocblas_status_success) return 2;
    const int lower = (ul == 'L') ? 1 : 0;
    const rocblas_fill fill = lower ? rocblas_fill_lower : rocblas_fill_upper;
    const int zcols = zcols_for(rg, n, il, iu);
    const size_t np = (size_t)n * (n + 1) / 2;
    const unsigned grid = grid_for((long long)n * n);
    Buf Bf(st);
    T* dA = Bf.get<T>((size_t)n * n);
    T* dB = Bf.get<T>((size_t)n * n);
    double* dW = Bf.get<double>(n);
    T* dZ = Bf.get<T>((size_t)n * zcols);
    T* dT = (jz == 'V' && it == 3) ? Bf.get<T>((size_t)n * zcols) : nullptr;
    T* dP = packed ? Bf.get<T>(np) : nullptr;
    rocblas_int* dinfo = Bf.get<rocblas_int>(1);
    if (!dA || !dB || !dW || !dZ || !dinfo || (jz == 'V' && it == 3 && !dT) || (packed && !dP)) return 2;
    if (packed) {
        if (hipMemcpyAsync(dP, A, sizeof(T) * np, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
        hipLaunchKernelGGL(k_dense_packed<T>, dim3(grid), dim3(256), 0, st, dP, dA, n, lower, 0);
        if (hipMemcpyAsync(dP, Bm, sizeof(T) * np, hipMemcpyHostToDevice, st) != hipSuccess) return 2;
        hipLaunchKernelGGL(k_dense_packed<T>, dim3(grid), dim3(256), 0, st, dP, dB, n, lower, 0);
    } else if (hipMemcpy2DAsync(dA, sizeof(T) * (size_t)n, A, sizeof(T) * (size_t)lda, sizeof(T) * (size_t)n,
                                n, hipMemcpyHostToDevice, st) != hipSuccess ||
               hipMemcpy2DAsync(dB, sizeof(T) * (size_t)n, Bm, sizeof(T) * (size_t)ldb, sizeof(T) * (size_t)n,
                                n, hipMemcpyHostToDevice, st) != hipSuccess) {
        return 2;
    }
    // Cholesky factor of B (dense.c:268)
    rocblas_int hinfo = 0;
    if (Lin<T>::potrf(c.h, fill, n, dB, n, dinfo) != rocblas_status_success ||
        hipMemcpyAsync(&hinfo, dinfo, sizeof(hinfo), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return 2;
    if (hinfo != 0) return 2;   // B is not positive definite (the reference asserts)
    // standard form (dense.c:272)
    const rocblas_eform ef = it == 1 ? rocblas_eform_ax : (it == 2 ? rocblas_eform_abx : rocblas_eform_bax);
    if (Lin<T>::gst(c.h, ef, fill, n, dA, n, dB, n) != rocblas_status_success) return 2;
    const int rc = standard_device<T>(jz, rg, ul, n, dA, n, vl, vu, il, iu, m, dW, dZ, n, zcols, st, c);
    if (rc) return rc;
    // back-transformation (dense.c:280-302): x = inv(L)^H y or inv(U) y; x = L y or U^H y
    T* dX = dZ;
    if (jz == 'V' && *m > 0) {
        if (it == 1 || it == 2) {
            const rocblas_operation tr = (ul == 'U') ? rocblas_operation_none : Lin<T>::adj;
            if (Lin<T>::trsm(c.h, fill, tr, n, *m, dB, n, dZ, n) != rocblas_status_success) return 2;
        } else {
            const rocblas_operation tr = (ul == 'U') ? Lin<T>::adj : rocblas_operation_none;
            if (Lin<T>::trmm(c.h, fill, tr, n, *m, dB, n, dZ, n, dT, n) != rocblas_status_success) return 2;
            dX = dT;
        }
    }
    if (*m > 0 && jz != 'C') {
        if (hipMemcpyAsync(W, dW, sizeof(double) * *m, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
        if (jz == 'V') {
            // eigenvectors into A (dense.c:307-308) or into Z (the packed drivers)
            T* dst = packed ? Z : A;
            const int ld = packed ? ldz : lda;
            if (copy_block_out(dst, ld, dX, n, *m, st) != 0) return 2;
        }
    }
    // B holds its Cholesky factor on return, as in the reference
    if (packed) {
        hipLaunchKernelGGL(k_dense_packed<T>, dim3(grid), dim3(256), 0, st, dP, dB, n, lower, 1);
        if (hipMemcpyAsync(Bm, dP, sizeof(T) * np, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
    } else if (hipMemcpy2DAsync(Bm, sizeof(T) * (size_t)ldb, dB, sizeof(T) * (size_t)n, sizeof(T) * (size_t)n,
                                n, hipMemcpyDeviceToHost, st) != hipSuccess) {
        return 2;
    }
    return hipStreamSynchronize(st) == hipSuccess ? 0 : 2;
}

// A leading dimension the caller gave: 0 stands for packed storage inside this file, so a
// non-positive one is mapped to -1 (invalid).
inline int ld_arg(const int* p) { return (p && *p > 0) ? *p : -1; }

}  // namespace

// ---------------------------------------------------------------- real symmetric

extern "C" int mr3_dsyeig_device(char jobz, char range, char uplo, int n, double* dA, int lda, double* vl,
                                 double* vu, int* il, int* iu, int* m, double* dW, double* dZ, int ldz,
                                 int zcols, void* stream) {
    const int jz = up(jobz), rg = up(range), ul = up(uplo);
    *m = 0;
    if (check_std(jz, rg, ul, n, vl, vu, il, iu) || lda < n) return 1;
    if (jz == 'V' && (dZ == nullptr || ldz < n || zcols < zcols_for(rg, n, il, iu))) return 1;
    Ctx& c = ctx();
    if (!c.ok) return 2;
    return standard_device<double>(jz, rg, ul, n, dA, lda, vl, vu, il, iu, m, dW, dZ, ldz, zcols,
                                   stream ? (hipStream_t)stream : c.st, c);
}

// dsyeig (mrrr.h:270-272, dense.c:151-230)
extern "C" int dsyeig(char* jobz, char* range, char* uplo, int* np, double* A, int* ldap, double* vl,
                      double* vu, int* il, int* iu, int* m, double* W, double* Z, int* ldz) {
    return standard_host<double>(up(jobz[0]), up(range[0]), up(uplo[0]), *np, A, ld_arg(ldap), vl, vu, il,
                                 iu, m, W, Z, ld_arg(ldz));
}

// dspeig (mrrr.h:444-446, dense.c:319-398): AP packed by columns.
extern "C" int dspeig(char* jobz, char* range, char* uplo, int* np, double* AP, double* vl, double* vu,
                      int* il, int* iu, int* m, double* W, double* Z, int* ldz) {
    return standard_host<double>(up(jobz[0]), up(range[0]), up(uplo[0]), *np, AP, 0, vl, vu, il, iu, m, W,
                                 Z, ld_arg(ldz));
}

// dsygeig (mrrr.h:352-356, dense.c:236-316): A x = l B x (1), A B x = l x (2), B A x = l x (3).
extern "C" int dsygeig(int* itype, char* jobz, char* range, char* uplo, int* np, double* A, int* ldap,
                       double* B, int* ldbp, double* vl, double* vu, int* il, int* iu, int* m, double* W) {
    if (*np <= 1) { *m = 0; return 1; }   // dense.c:256
    return general_host<double>(*itype, up(jobz[0]), up(range[0]), up(uplo[0]), *np, A, ld_arg(ldap), B,
                                ld_arg(ldbp), vl, vu, il, iu, m, W, nullptr, 0);
}

// dspgeig (mrrr.h:526-530, dense.c:403-471)
extern "C" int dspgeig(int* itype, char* jobz, char* range, char* uplo, int* np, double* AP, double* BP,
                       double* vl, double* vu, int* il, int* iu, int* m, double* W, double* Z, int* ldz) {
    if (*np < 1) { *m = 0; return 1; }    // dense.c:422
    return general_host<double>(*itype, up(jobz[0]), up(range[0]), up(uplo[0]), *np, AP, 0, BP, 0, vl, vu,
                                il, iu, m, W, Z, ld_arg(ldz));
}

// --------------------------------------------------------------- complex Hermitian

// zheeig (mrrr.h:538-541, dense.c:568-672)
extern "C" int zheeig(char* jobz, char* range, char* uplo, int* np, void* A, int* ldap, double* vl,
                      double* vu, int* il, int* iu, int* m, double* W, void* Z, int* ldz) {
    return standard_host<zc>(up(jobz[0]), up(range[0]), up(uplo[0]), *np, (const zc*)A, ld_arg(ldap), vl, vu,
                             il, iu, m, W, (zc*)Z, ld_arg(ldz));
}

// zhpeig (mrrr.h:718-721, dense.c:760-865)
extern "C" int zhpeig(char* jobz, char* range, char* uplo, int* np, void* AP, double* vl, double* vu,
                      int* il, int* iu, int* m, double* W, void* Z, int* ldz) {
    return standard_host<zc>(up(jobz[0]), up(range[0]), up(uplo[0]), *np, (const zc*)AP, 0, vl, vu, il, iu,
                             m, W, (zc*)Z, ld_arg(ldz));
}

// zhegeig (mrrr.h:624-628, dense.c:676-757)
extern "C" int zhegeig(int* itype, char* jobz, char* range, char* uplo, int* np, void* A, int* ldap,
                       void* B, int* ldbp, double* vl, double* vu, int* il, int* iu, int* m, double* W) {
    if (*np <= 1) { *m = 0; return 1; }   // dense.c:697
    return general_host<zc>(*itype, up(jobz[0]), up(range[0]), up(uplo[0]), *np, (zc*)A, ld_arg(ldap),
                            (zc*)B, ld_arg(ldbp), vl, vu, il, iu, m, W, nullptr, 0);
}

// zhpgeig (mrrr.h:802-806, dense.c:869-935)
extern "C" int zhpgeig(int* itype, char* jobz, char* range, char* uplo, int* np, void* AP, void* BP,
                       double* vl, double* vu, int* il, int* iu, int* m, double* W, void* Z, int* ldz) {
    if (*np <= 1) { *m = 0; return 1; }   // dense.c:888
    return general_host<zc>(*itype, up(jobz[0]), up(range[0]), up(uplo[0]), *np, (zc*)AP, 0, (zc*)BP, 0, vl,
                            vu, il, iu, m, W, (zc*)Z, ld_arg(ldz));
}

// ------------------------------------------------- Fortran bindings (dense.c:1037-1114)

extern "C" void dsyeig_(char* jobz, char* range, char* uplo, int* n, double* A, int* lda, double* vl,
                        double* vu, int* il, int* iu, int* m, double* W, double* Z, int* ldz, int* info) {
    *info = dsyeig(jobz, range, uplo, n, A, lda, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void dsygeig_(int* itype, char* jobz, char* range, char* uplo, int* n, double* A, int* lda,
                         double* B, int* ldb, double* vl, double* vu, int* il, int* iu, int* m, double* W,
                         int* info) {
    *info = dsygeig(itype, jobz, range, uplo, n, A, lda, B, ldb, vl, vu, il, iu, m, W);
}

extern "C" void dspeig_(char* jobz, char* range, char* uplo, int* n, double* AP, double* vl, double* vu,
                        int* il, int* iu, int* m, double* W, double* Z, int* ldz, int* info) {
    *info = dspeig(jobz, range, uplo, n, AP, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void dspgeig_(int* itype, char* jobz, char* range, char* uplo, int* n, double* AP, double* BP,
                         double* vl, double* vu, int* il, int* iu, int* m, double* W, double* Z, int* ldz,
                         int* info) {
    *info = dspgeig(itype, jobz, range, uplo, n, AP, BP, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void zheeig_(char* jobz, char* range, char* uplo, int* n, void* A, int* lda, double* vl,
                        double* vu, int* il, int* iu, int* m, double* W, void* Z, int* ldz, int* info) {
    *info = zheeig(jobz, range, uplo, n, A, lda, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void zhegeig_(int* itype, char* jobz, char* range, char* uplo, int* n, void* A, int* lda,
                         void* B, int* ldb, double* vl, double* vu, int* il, int* iu, int* m, double* W,
                         int* info) {
    *info = zhegeig(itype, jobz, range, uplo, n, A, lda, B, ldb, vl, vu, il, iu, m, W);
}

extern "C" void zhpeig_(char* jobz, char* range, char* uplo, int* n, void* AP, double* vl, double* vu,
                        int* il, int* iu, int* m, double* W, void* Z, int* ldz, int* info) {
    *info = zhpeig(jobz, range, uplo, n, AP, vl, vu, il, iu, m, W, Z, ldz);
}

extern "C" void zhpgeig_(int* itype, char* jobz, char* range, char* uplo, int* n, void* AP, void* BP,
                         double* vl, double* vu, int* il, int* iu, int* m, double* W, void* Z, int* ldz,
                         int* info) {
    *info = zhpgeig(itype, jobz, range, uplo, n, AP, BP, vl, vu, il, iu, m, W, Z, ldz);
}
