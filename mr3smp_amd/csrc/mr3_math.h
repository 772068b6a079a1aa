// mr3_math.h -- per-lane numerical kernels of the MRRR algorithm (device code).
//
// Every routine is a single-lane restatement of one vendored LAPACK kernel of the
// reference; the GPU parallelism comes from running one lane per eigenvalue
// interval or per eigenvector.  All arithmetic is binary64 and the library is built
// with -ffp-contract=off, so no multiply-add is fused and the Sturm counts carry
// the same roundings as the reference's x86-64 build.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

namespace mr3 {

constexpr double kEps = DBL_EPSILON;   // odmch.f 'P'
constexpr double kSafmin = DBL_MIN;    // odmch.f 'S' / 'U'
constexpr int kMaxExpand = 128;        // cap on bracket doublings (2^128 spans any range)

// Loop chunks: every sweep runs in chunks of kU steps and issues the loads of the
// next chunk before it computes the current one (software pipelining), so a chunk's
// memory latency is covered by the previous chunk's arithmetic instead of being paid
// by every dependent step.  The prefetch reads up to 2 kU entries past either end of
// the arrays unclamped: every array a sweep reads is allocated with kSweepPad spare
// entries on both sides, and entries read outside the block are never used.
constexpr int kU = 8;

// Read-only device arrays that every active lane of a wave reads at the same index
// can go through the constant address space: the compiler then loads them with
// scalar loads into SGPRs (no per-lane address arithmetic, one fetch per wave) and
// the FP64 operations read them from there.  The sweeps below are templates over the
// pointer type P: `const double*` (per-lane arrays) or CPtr (wave-uniform arrays,
// chosen at run time by the caller with wave_uniform()).
typedef const __attribute__((address_space(4))) double* CPtr;

__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ const double* rfl(const double* p) {
    const unsigned long long u = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(u & 0xffffffffu));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(u >> 32));
    return (const double*)(((unsigned long long)hi << 32) | lo);
}
// Arrays that the lanes of a wave read at different addresses go through the global
// address space.  A load through a generic pointer (flat_load) also counts against the
// LDS / scalar-memory counter, so the compiler's wait for any LDS or scalar result
// would also wait for the loads a sweep has prefetched for its next chunk.
typedef const __attribute__((address_space(1))) double* GPtr;
typedef __attribute__((address_space(1))) double* GPtrW;
typedef __attribute__((address_space(1))) double GDouble;
typedef __attribute__((address_space(1))) int* GIntW;
__device__ __forceinline__ GPtr as_global(const double* p) { return (GPtr)p; }
__device__ __forceinline__ GPtrW as_global_w(double* p) { return (GPtrW)p; }
__device__ __forceinline__ GIntW as_global_w(int* p) { return (GIntW)p; }

// True when every active lane holds the same value (wave-uniform result).
__device__ __forceinline__ bool wave_uniform(int v) { return __all(v == rfl(v)); }
__device__ __forceinline__ bool wave_uniform(const double* p) { return __all(p == rfl(p)); }
__device__ __forceinline__ CPtr as_const(const double* p) { return (CPtr)rfl(p); }

// Sturm count of T - x I, dlaebz IJOB=1 convention (odebz.f:265-277): a pivot with
// |t| < pivmin becomes -pivmin, a pivot <= 0 counts.
template <class P>
__device__ __forceinline__ int count_T_ijob1(int n, P d, P e2, double x, double pivmin) {
    double t = d[0] - x;
    if (fabs(t) < pivmin) t = -pivmin;
    int c = (t <= 0.0);
    if (n < 2) return c;
    double dd[kU], ee[kU], dn[kU], en[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        int j = 1 + u;
        dd[u] = d[j]; ee[u] = e2[j - 1];
    }
    for (int j0 = 1; j0 < n; j0 += kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int j = j0 + kU + u;
            dn[u] = d[j]; en[u] = e2[j - 1];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (j0 + u < n) {
                t = dd[u] - ee[u] / t - x;
                if (fabs(t) < pivmin) t = -pivmin;
                c += (t <= 0.0);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) { dd[u] = dn[u]; ee[u] = en[u]; }
    }
    return c;
}

// Sturm count of T - x I, dlaebz IJOB=2/3 convention (odebz.f:395-415): a pivot
// <= pivmin counts and is replaced by min(t, -pivmin).
template <class P>
__device__ __forceinline__ int count_T_ijob2(int n, P d, P e2, double x, double pivmin) {
    double t = d[0] - x;
    int c = 0;
    if (t <= pivmin) { c = 1; t = fmin(t, -pivmin); }
    if (n < 2) return c;
    double dd[kU], ee[kU], dn[kU], en[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        int j = 1 + u;
        dd[u] = d[j]; ee[u] = e2[j - 1];
    }
    for (int j0 = 1; j0 < n; j0 += kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int j = j0 + kU + u;
            dn[u] = d[j]; en[u] = e2[j - 1];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (j0 + u < n) {
                t = dd[u] - ee[u] / t - x;
                if (t <= pivmin) { ++c; t = fmin(t, -pivmin); }
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) { dd[u] = dn[u]; ee[u] = en[u]; }
    }
    return c;
}

// dlaneg (odneg.f:97-164): negcount of L D L^T - sigma I at twist r (1-based).
// The reference runs blocks of 128 steps and re-runs a block with Inf/Inf -> 1
// when the block ends in NaN.  A NaN in T can only come from the quotient and then
// persists to the end of the block, and the re-run differs from the fast loop only
// at quotients that are NaN; substituting per step therefore gives the same count.
template <class P>
__device__ __forceinline__ int dlaneg(int n, P d, P lld, double sigma, int r) {
    int neg = 0;
    double t = -sigma;
    const int top = r - 1;   // steps j = 0 .. r-2
    double dd[kU], ll[kU], dn[kU], ln[kU];
    if (top > 0) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int j = u;
            dd[u] = d[j]; ll[u] = lld[j];
        }
        for (int j0 = 0; j0 < top; j0 += kU) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                int j = j0 + kU + u;
                dn[u] = d[j]; ln[u] = lld[j];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (j0 + u < top) {
                    double dplus = dd[u] + t;
                    neg += (dplus < 0.0);
                    double tmp = t / dplus;
                    tmp = isnan(tmp) ? 1.0 : tmp;
                    t = tmp * ll[u] - sigma;
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) { dd[u] = dn[u]; ll[u] = ln[u]; }
        }
    }
    double p = d[n - 1] - sigma;
    if (n - 2 >= r - 1) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int j = n - 2 - u;
            dd[u] = d[j]; ll[u] = lld[j];
        }
        for (int j0 = n - 2; j0 >= r - 1; j0 -= kU) {
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                int j = j0 - kU - u;
                dn[u] = d[j]; ln[u] = lld[j];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                if (j0 - u >= r - 1) {
                    double dminus = ll[u] + p;
                    neg += (dminus < 0.0);
                    double tmp = p / dminus;
                    tmp = isnan(tmp) ? 1.0 : tmp;
                    p = tmp * dd[u] - sigma;
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) { dd[u] = dn[u]; ll[u] = ln[u]; }
        }
    }
    double gamma = (t + sigma) + p;
    neg += (gamma < 0.0);
    return neg;
}

// dlarrb (odrrb.f:135-299) for one interval.  The reference keeps a linked list of
// intervals but never lets them interact: each is bisected against its own index
// with a gap taken from the (unchanging) WGAP input, so one lane per interval is
// exact.  `gap` is min(left gap, right gap) as dlarrb forms it.  On return *w and
// *werr hold the refined midpoint and half width, or are left unchanged when the
// interval already met the tolerance (odrrb.f:198-205).
template <class P>
__device__ __forceinline__ int dlarrb1(int n, P d, P lld, int i, double rtol1, double rtol2,
                                        double gap, double* w, double* werr, double pivmin,
                                        double spdiam, int twist) {
    const int maxitr = (int)((log(spdiam + pivmin) - log(pivmin)) / log(2.0)) + 2;
    const double mnwdth = 2.0 * pivmin;
    int r = twist;
    if (r < 1 || r > n) r = n;
    double left = *w - *werr, right = *w + *werr;
    int sweeps = 0;
    // The reference expands without a bound (odrrb.f:176-196); a zero error bound
    // would never move the end point, so the expansion is capped here.
    double back = *werr;
    for (int g = 0; g < kMaxExpand && (++sweeps, dlaneg(n, d, lld, left, r) > i - 1); ++g) {
        left = left - back;
        back = 2.0 * back;
    }
    back = *werr;
    for (int g = 0; g < kMaxExpand && (++sweeps, dlaneg(n, d, lld, right, r) < i); ++g) {
        right = right + back;
        back = 2.0 * back;
    }
    double width = 0.5 * fabs(left - right);
    double tmp = fmax(fabs(left), fabs(right));
    double cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
    if (width <= cvrgd || width <= mnwdth) return sweeps;
    for (int iter = 0;; ++iter) {
        double mid = 0.5 * (left + right);
        width = right - mid;
        tmp = fmax(fabs(left), fabs(right));
        cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
        if (width <= cvrgd || width <= mnwdth || iter == maxitr) break;
        int negcnt = dlaneg(n, d, lld, mid, r);
        ++sweeps;
        if (negcnt <= i - 1) left = mid;
        else right = mid;
    }
    *w = 0.5 * (left + right);
    *werr = right - *w;
    return sweeps;
}

// Sturm count on T for dlarrj (odrrj.f:160-166): no pivmin clamp, count if < 0.
template <class P>
__device__ __forceinline__ int count_T_rrj(int n, P d, P e2, double s) {
    double dplus = d[0] - s;
    int cnt = (dplus < 0.0);
    if (n < 2) return cnt;
    double dd[kU], ee[kU], dn[kU], en[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        int j = 1 + u;
        dd[u] = d[j]; ee[u] = e2[j - 1];
    }
    for (int j0 = 1; j0 < n; j0 += kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            int j = j0 + kU + u;
            dn[u] = d[j]; en[u] = e2[j - 1];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (j0 + u < n) {
                dplus = dd[u] - s - ee[u] / dplus;
                cnt += (dplus < 0.0);
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) { dd[u] = dn[u]; ee[u] = en[u]; }
    }
    return cnt;
}

// dlarrj (odrrj.f:100-281) for one interval (intervals are independent).
template <class P>
__device__ __forceinline__ int dlarrj1(int n, P d, P e2, int i, double rtol, double* w,
                                        double* werr, double pivmin, double spdiam) {
    const int maxitr = (int)((log(spdiam + pivmin) - log(pivmin)) / log(2.0)) + 2;
    double left = *w - *werr, mid = *w, right = *w + *werr;
    double width = right - mid;
    double tmp = fmax(fabs(left), fabs(right));
    if (width < rtol * tmp) return 0;
    int sweeps = 0;
    double fac = 1.0;
    for (int g = 0; g < kMaxExpand && (++sweeps, count_T_rrj(n, d, e2, left) > i - 1); ++g) {
        left = left - *werr * fac;
        fac = 2.0 * fac;
    }
    fac = 1.0;
    for (int g = 0; g < kMaxExpand && (++sweeps, count_T_rrj(n, d, e2, right) < i); ++g) {
        right = right + *werr * fac;
        fac = 2.0 * fac;
    }
    for (int iter = 0;; ++iter) {
        mid = 0.5 * (left + right);
        width = right - mid;
        tmp = fmax(fabs(left), fabs(right));
        if (width < rtol * tmp || iter == maxitr) break;
        int cnt = count_T_rrj(n, d, e2, mid);
        ++sweeps;
        if (cnt <= i - 1) left = mid;
        else right = mid;
    }
    *w = 0.5 * (left + right);
    *werr = right - *w;
    return sweeps;
}

// 48-bit LCG of dlaruv (odruv.f:330-387): x_i = seed * a^i mod 2^48.  dlarnv with
// IDIST=2 (odrnv.f:60-116) consumes the stream in chunks of 64 and hands the last
// value of a chunk on as the next seed, so entry i (0-based) of the vector is
// 2 * (seed * a^(i+1) mod 2^48) / 2^48 - 1.
constexpr unsigned long long kLcgA = 33952834046453ULL;
constexpr unsigned long long kMask48 = (1ULL << 48) - 1ULL;

__host__ __device__ __forceinline__ unsigned long long lcg_pow(unsigned long long e) {
    unsigned long long r = 1ULL, b = kLcgA;
    while (e) {
        if (e & 1ULL) r = (r * b) & kMask48;
        b = (b * b) & kMask48;
        e >>= 1;
    }
    return r;
}

__host__ __device__ __forceinline__ double lcg_uniform_pm1(unsigned long long x) {
    const double R = 1.0 / 4096.0;
    double it1 = (double)((x >> 36) & 4095ULL), it2 = (double)((x >> 24) & 4095ULL);
    double it3 = (double)((x >> 12) & 4095ULL), it4 = (double)(x & 4095ULL);
    double u = R * (it1 + R * (it2 + R * (it3 + R * it4)));
    return 2.0 * u - 1.0;
}

// Dispatchers: the scalar-load form when every active lane of the wave sweeps the
// same arrays with the same length and twist (the usual case: consecutive
// eigenvalues of one block or cluster), the per-lane form otherwise.  Both forms run
// the same operations, so results do not depend on the choice.
__device__ __forceinline__ int count_T_ijob2_w(int n, const double* d, const double* e2,
                                               double x, double pivmin) {
    if (wave_uniform(d) && wave_uniform(e2) && wave_uniform(n))
        return count_T_ijob2(rfl(n), as_const(d), as_const(e2), x, pivmin);
    return count_T_ijob2(n, d, e2, x, pivmin);
}

__device__ __forceinline__ int dlarrb1_w(int n, const double* d, const double* lld, int i,
                                         double rtol1, double rtol2, double gap, double* w,
                                         double* werr, double pivmin, double spdiam, int twist) {
    if (wave_uniform(d) && wave_uniform(lld) && wave_uniform(n) && wave_uniform(twist))
        return dlarrb1(rfl(n), as_const(d), as_const(lld), i, rtol1, rtol2, gap, w, werr, pivmin,
                       spdiam, rfl(twist));
    return dlarrb1(n, as_global(d), as_global(lld), i, rtol1, rtol2, gap, w, werr, pivmin, spdiam,
                   twist);
}

__device__ __forceinline__ int dlarrj1_w(int n, const double* d, const double* e2, int i,
                                         double rtol, double* w, double* werr, double pivmin,
                                         double spdiam) {
    if (wave_uniform(d) && wave_uniform(e2) && wave_uniform(n))
        return dlarrj1(rfl(n), as_const(d), as_const(e2), i, rtol, w, werr, pivmin, spdiam);
    return dlarrj1(n, as_global(d), as_global(e2), i, rtol, w, werr, pivmin, spdiam);
}

}  // namespace mr3
