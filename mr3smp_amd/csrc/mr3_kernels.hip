// mr3_kernels.hip -- HIP kernels of the MRRR tridiagonal eigensolver for gfx950.
//
// Stage map (reference file:line -> kernel):
//   mrrr.c:325-365 scale, odrrr.f dlarrr, mrrr_val.c:214-340 Gershgorin/split/counts
//                                                    -> k_prep (one workgroup per matrix)
//   mrrr_val.c:421 find_eigval_irreducible / dlarrd  -> k_coarse (one lane per eigenvalue)
//   mrrr_val.c:520-626 surplus discard (range I)     -> k_discard
//   mrrr_val.c:717-948 compute_root_rrr + dlarnv      -> k_root (one workgroup per block)
//   mrrr_val.c:954-1087 refine_eigval_approx / dlarrb -> k_refine_root + k_gaps_root
//   mrrr_vec.c:380-416 init_zindex                    -> k_zindex
//   process_cluster.c:240-433 refine (depth 0) and PMR_create_subtasks :438-629
//                                                    -> k_classify_root / k_classify
//   process_cluster.c:105-235 compute_new_rrr / dlarrf -> k_newrrr (one wave per cluster)
//   process_cluster.c:366-432 cluster refinement       -> k_refine_cl + k_gaps_cl
//   process_singleton.c:65-324 S-task / dlar1v         -> k_s_iter, k_s_bisect,
//                                                        k_s_collect, k_s_normalize
//   mrrr.c:263-288 shift back, dlarrj, unscale, sort   -> k_shift_n, k_rrj, k_finalize
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mr3_internal.h"
#include "mr3_math.h"

namespace mr3 {

// Algorithmic work counters, in sweep elements per kernel family; a lane adds its
// total once at the end, and only when the host enabled counting.
enum StepKind : int {
    kStepCoarse = 0,     // Sturm steps on T, 3 flops (odebz.f:410)
    kStepRefine = 1,     // dlaneg steps in the root refinement, 4 flops (odneg.f:106-109)
    kStepEndref = 2,     // dlaneg steps refining cluster ends, 4 flops
    kStepR1v = 3,        // dlar1v elements, 16 flops (odr1v.f:185-362)
    kStepR1vBis = 4,     // dlaneg steps of the singleton bisection fallback, 4 flops
    kStepRrf = 5,        // dlarrf elements per sweep, 5 flops (odrrf.f:212-214)
    kStepRrj = 6,        // dlarrj Sturm steps, 3 flops
    kStepRefineCl = 7,   // dlaneg steps in cluster refinement, 4 flops
    kStepR1vWave = 8,    // dlar1v elements executed per wave (lane 0 counts n per call)
    kStepBisWave = 9,    // bisection-fallback elements of the slowest lane of a wave
    kStepTaskTicks = 10, // wave time of the singleton tasks, 100 MHz wall-clock ticks
    kStepTasks = 11,     // singleton tasks run
    kStepTBis = 12,      // wave time (ticks) in the bisection fallback
    kStepTR1v = 13,      // wave time in the twisted factorisations (dlar1v to the twist)
    kStepTSolve = 14,    // wave time in the vector solves and column writes
    kStepRrfN = 15,      // block rows of the clusters given to dlarrf (one count per cluster)
    kStepSmDqds = 16,    // small-matrix route: dqds steps, 5 flops
    kStepSmNeg = 17,     //   dlaneg steps, 4 flops
    kStepSmSturm = 18,   //   Sturm steps on T, 3 flops
    kStepSmR1v = 19,     //   dlar1v rows, 16 flops
    kStepSmRrf = 20,     //   dlarrf elements, 5 flops
    kStepTailR1v = 21,   // dlar1v elements of the tail kernel (S and P in LDS), 16 flops
    kStepTailBis = 22,   //   its bisection-fallback dlaneg steps, 4 flops
    kStepTailBR1v = 23,  // the same for the large-block tail kernel (S and P in device memory)
    kStepTailBBis = 24,
    kStepFallbacks = 25, // singletons that fall back to bisection (process_singleton.c:186-204),
                         // one count per fallback
    kStepKinds = 26
};
__device__ unsigned long long g_steps[kStepKinds];
__constant__ int g_count_steps;

__device__ __forceinline__ void add_steps(int kind, unsigned long long v) {
    if (g_count_steps && v) atomicAdd(&g_steps[kind], v);
}

// Lanes per eigenvalue of the cluster refinements of a level with `items` eigenvalues:
// 1 (the one-lane kernels) unless a group of 4 or more lanes per eigenvalue keeps the
// launch within the lane budget; then the largest such power of two up to a wave.
// Multisection with G lanes does (G - 1) / log2(G) times the sweeps of one-point
// bisection in 1 / log2(G) of the sequential rounds.  The kernels of a level read
// their item counts on the device and pick their variant with it.
__host__ __device__ inline int group_lanes_hd(long long items, long long budget) {
    if (items * 4 > budget) return 1;
    int G = 4;
    while (G < 64 && items * 2 * G <= budget) G *= 2;
    return G;
}


// ---------------------------------------------------------------------------------
// Workgroup helpers (256 threads)
// ---------------------------------------------------------------------------------
template <int NT>
__device__ double wg_reduce_max(double v, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            double a = sh[tid], b = sh[tid + s];
            sh[tid] = (isnan(a) || a > b) ? a : b;
        }
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <int NT>
__device__ double wg_reduce_min(double v, double* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = fmin(sh[tid], sh[tid + s]);
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

template <int NT>
__device__ int wg_reduce_and(int v, int* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) sh[tid] = sh[tid] & sh[tid + s];
        __syncthreads();
    }
    int r = sh[0];
    __syncthreads();
    return r;
}

// Exclusive scan of one int per thread; returns the prefix, *total gets the sum.
template <int NT>
__device__ int wg_exscan(int v, int* sh, int* total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = 1; s < NT; s <<= 1) {
        int t = (tid >= s) ? sh[tid - s] : 0;
        __syncthreads();
        sh[tid] += t;
        __syncthreads();
    }
    int incl = sh[tid];
    *total = sh[NT - 1];
    __syncthreads();
    return incl - v;
}

// dlaev2 (odev2.f:70-170): eigen-decomposition of the 2x2 [[a,b],[b,c]].
__device__ void dlaev2(double a, double b, double c, double* rt1, double* rt2, double* cs1,
                       double* sn1) {
    double sm = a + c, df = a - c, adf = fabs(df), tb = b + b, ab = fabs(tb);
    double acmx, acmn, rt, cs, ct, tn, acs;
    int sgn1, sgn2;
    if (fabs(a) > fabs(c)) { acmx = a; acmn = c; } else { acmx = c; acmn = a; }
    if (adf > ab) rt = adf * sqrt(1.0 + (ab / adf) * (ab / adf));
    else if (adf < ab) rt = ab * sqrt(1.0 + (adf / ab) * (adf / ab));
    else rt = ab * sqrt(2.0);
    if (sm < 0.0) { *rt1 = 0.5 * (sm - rt); sgn1 = -1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else if (sm > 0.0) { *rt1 = 0.5 * (sm + rt); sgn1 = 1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else { *rt1 = 0.5 * rt; *rt2 = -0.5 * rt; sgn1 = 1; }
    if (df >= 0.0) { cs = df + rt; sgn2 = 1; } else { cs = df - rt; sgn2 = -1; }
    acs = fabs(cs);
    if (acs > ab) { ct = -tb / cs; *sn1 = 1.0 / sqrt(1.0 + ct * ct); *cs1 = ct * *sn1; }
    else if (ab == 0.0) { *cs1 = 1.0; *sn1 = 0.0; }
    else { tn = -cs / tb; *cs1 = 1.0 / sqrt(1.0 + tn * tn); *sn1 = tn * *cs1; }
    if (sgn1 == sgn2) { tn = *cs1; *cs1 = -*sn1; *sn1 = tn; }
}

// dstemr's closed forms for n = 1 and n = 2 (odstmr.f:357-430); mrrr() reaches them
// through mrrr.c:149-153.  Results are final (unscaled, sorted as dstemr leaves them).
__device__ void small_case(const MatDesc& d, MatState& S, const double* D, const double* E) {
    int wantz = (d.jobz == kV);
    int m = 0;
    if (d.n == 1) {
        if (d.range == kAll || d.range == kInd) { m = 1; d.Wout[0] = D[0]; }
        else if (d.vl < D[0] && d.vu >= D[0]) { m = 1; d.Wout[0] = D[0]; }
        if (wantz && m > 0) { d.Z[0] = 1.0; d.Zsupp[0] = 1; d.Zsupp[1] = 1; }
    } else {
        double r1, r2, cs, sn;
        dlaev2(D[0], E[0], D[1], &r1, &r2, &cs, &sn);
        if (d.range == kAll || (d.range == kVal && r2 > d.vl && r2 <= d.vu) ||
            (d.range == kInd && d.il == 1)) {
            ++m;
            d.Wout[m - 1] = r2;
            if (wantz) {
                double* z = d.Z + (size_t)(m - 1) * d.ldz;
                z[0] = -sn;
                z[1] = cs;
                int* zs = d.Zsupp + 2 * (m - 1);
                if (sn != 0.0) { zs[0] = 1; zs[1] = (cs != 0.0) ? 2 : 1; }
                else { zs[0] = 2; zs[1] = 2; }
            }
        }
        if (d.range == kAll || (d.range == kVal && r1 > d.vl && r1 <= d.vu) ||
            (d.range == kInd && d.iu == 2)) {
            ++m;
            d.Wout[m - 1] = r1;
            if (wantz) {
                double* z = d.Z + (size_t)(m - 1) * d.ldz;
                z[0] = cs;
                z[1] = sn;
                int* zs = d.Zsupp + 2 * (m - 1);
                if (sn != 0.0) { zs[0] = 1; zs[1] = (cs != 0.0) ? 2 : 1; }
                else { zs[0] = 2; zs[1] = 2; }
            }
        }
    }
    S.m = m;
    S.mtot = m;
    S.vl = d.vl;
    S.vu = d.vu;
    S.il = d.il;
    S.iu = d.iu;
    S.tryrac = d.tryrac;
    S.done = 1;
    S.status = kOk;
}

// dlaebz IJOB=3 for the two target counts il - 1 and iu (mrrr_val.c:295-319, the
// loop of odebz.f:380-536 with one interval).  Each target has half of the workgroup
// (the two intervals never interact).  A bisection point depends only on the interval
// and the interval only on the decisions before it, so the points of the next kL3
// steps of every possible path form a binary tree below the current point (node t:
// root 1, child 2t when the count exceeds the target, 2t + 1 when it is below).  The
// half evaluates all of them at once, one count per lane, then its first lane walks
// the path the sequential loop takes, with its updates and its stopping test step by
// step.  Points, counts and results are the sequential loop's bit for bit, with about
// a seventh of its sweeps one after the other.
// The block's fudged Gershgorin interval as dlarrd's block loop forms it
// (odrrd.f:337-341 with FUDGE = 2, the block's own norm).
__device__ __forceinline__ void coarse_bounds(const BlockRec& br, int nbl, double pivmin,
                                              double* left, double* right) {
    const double tnormb = fmax(fabs(br.glraw), fabs(br.guraw));
    *left = br.glraw - 2.0 * tnormb * kEps * nbl - 2.0 * pivmin;
    *right = br.guraw + 2.0 * tnormb * kEps * nbl + 2.0 * pivmin;
}

constexpr int PREP_NT = 256;   // threads of k_prep
constexpr int kL3 = 7;
struct Ijob3 {
    double a1, a2, c;
    int n1, n2, it, done;
};

__device__ void ijob3_pair(int n, const double* d, const double* e2, double pivmin,
                           double abstol, double reltol, int nitmax, const int nval[2],
                           Ijob3* st, int (*cnt)[1 << kL3]) {
    const int half = threadIdx.x / (PREP_NT / 2), g = threadIdx.x % (PREP_NT / 2);
    static_assert(PREP_NT / 2 >= (1 << kL3) - 1, "one lane per tree node");
    for (;;) {
        __syncthreads();
        if (st[0].done && st[1].done) break;
        Ijob3& q = st[half];
        const int t = g + 1;
        if (!q.done && t < (1 << kL3)) {
            const int dt = 31 - __clz(t);
            if (q.it + dt < nitmax) {
                double a1 = q.a1, a2 = q.a2, c = q.c;
                for (int b = dt - 1; b >= 0; --b) {
                    if ((t >> b) & 1) a1 = c;
                    else a2 = c;
                    c = 0.5 * (a1 + a2);
                }
                cnt[half][t] = count_T_ijob2_w(n, d, e2, c, pivmin);
            }
        }
        __syncthreads();
        if (g == 0 && !q.done) {
            const int nv = nval[half];
            double a1 = q.a1, a2 = q.a2, c = q.c;
            int n1 = q.n1, n2 = q.n2, it = q.it, done = 0, t = 1;
            for (int lv = 0; lv < kL3; ++lv) {
                if (it >= nitmax) { done = 1; break; }
                const int k = cnt[half][t];
                if (k <= nv) { a1 = c; n1 = k; }
                if (k >= nv) { a2 = c; n2 = k; }
                ++it;
                const double tmp1 = fabs(a2 - a1), tmp2 = fmax(fabs(a2), fabs(a1));
                if (tmp1 < fmax(fmax(abstol, pivmin), reltol * tmp2) || n1 >= n2) { done = 1; break; }
                c = 0.5 * (a1 + a2);
                t = 2 * t + (k < nv ? 1 : 0);
            }
            if (it >= nitmax) done = 1;
            q.a1 = a1; q.a2 = a2; q.c = c; q.n1 = n1; q.n2 = n2; q.it = it; q.done = done;
        }
    }
}

// ---------------------------------------------------------------------------------
// k_prep: scaling, rel-acc test, Gershgorin, pivmin, splitting, counts and the
// per-block index ranges of one matrix (mrrr.c:155-243, mrrr_val.c:110-438).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(PREP_NT) void k_prep(const MatDesc* __restrict__ md,
                                                  MatState* __restrict__ ms, double* D, double* E,
                                                  double* E2, double* gers, double* Dcopy,
                                                  double* E2copy, int* isplit, BlockRec* blk,
                                                  double* W, double* Werr, double* Wcoarse,
                                                  int* iblock, int* Windex, int* smat) {
    __shared__ double shd[PREP_NT];
    __shared__ int shi[PREP_NT];
    __shared__ int sh_c0, sh_c1;
    const int k = blockIdx.x, tid = threadIdx.x;
    const MatDesc dsc = md[k];
    MatState& S = ms[k];
    const int n = dsc.n, off = dsc.off;
    double* d = D + off;
    double* e = E + off;
    double* e2 = E2 + off;
    double* g = gers + 2 * (size_t)off;

    for (int i = tid; i < n; i += PREP_NT) {
        d[i] = dsc.Din[i];
        e[i] = (i < n - 1) ? dsc.Ein[i] : 0.0;
        smat[off + i] = k;
    }
    if (tid == 0) {
        S.status = kOk;
        S.done = 0;
        S.m = 0;
        S.mtot = 0;
        S.nsplit = 0;
        S.maxbl = 0;
        S.scale = 1.0;
        S.vl = dsc.vl;
        S.vu = dsc.vu;
        S.il = dsc.il;
        S.iu = dsc.iu;
        S.tryrac = dsc.tryrac;
    }
    __syncthreads();
    if (n <= 2 && dsc.jobz != kC) {
        if (tid == 0) small_case(dsc, S, d, e);
        return;
    }
    // ---- scale_matrix (mrrr.c:325-365)
    double tmax = 0.0;
    for (int i = tid; i < n; i += PREP_NT) {
        double a = fabs(d[i]);
        if (isnan(a) || a > tmax) tmax = a;
        if (i < n - 1) {
            a = fabs(e[i]);
            if (isnan(a) || a > tmax) tmax = a;
        }
    }
    double T_norm = wg_reduce_max<PREP_NT>(tmax, shd);
    const double smlnum = DBL_MIN / DBL_EPSILON, bignum = 1.0 / smlnum;
    const double rmin = sqrt(smlnum), rmax = fmin(sqrt(bignum), 1.0 / sqrt(sqrt(DBL_MIN)));
    double scale = 1.0;
    if (T_norm > 0 && T_norm < rmin) scale = rmin / T_norm;
    else if (T_norm > rmax) scale = rmax / T_norm;
    double vl = dsc.vl, vu = dsc.vu;
    if (scale != 1.0) {
        for (int i = tid; i < n; i += PREP_NT) {
            d[i] *= scale;
            if (i < n - 1) e[i] *= scale;
        }
        if (dsc.range == kVal) { vl *= scale; vu *= scale; }
        __syncthreads();
    }
    // ---- dlarrr (odrrr.f:60-121), evaluated for all i at once
    int tryrac = dsc.tryrac;
    int relok = 0;
    if (tryrac) {
        const double rmin2 = sqrt(DBL_MIN / DBL_EPSILON);
        int ok = 1;
        for (int i = tid; i < n; i += PREP_NT) {
            double ti = sqrt(fabs(d[i]));
            if (ti < rmin2) ok = 0;
            if (i >= 1) {
                double tp = sqrt(fabs(d[i - 1]));
                double od = fabs(e[i - 1]) / (tp * ti);
                double odp = 0.0;
                if (i >= 2) odp = fabs(e[i - 2]) / (sqrt(fabs(d[i - 2])) * tp);
                if (!(tp < rmin2) && !(ti < rmin2) && odp + od >= 0.999) ok = 0;
            }
        }
        relok = wg_reduce_and<PREP_NT>(ok, shi);
    }
    double split;
    if (tryrac && relok) {
        split = DBL_EPSILON;
        for (int i = tid; i < n; i += PREP_NT) {
            Dcopy[off + i] = d[i];
            E2copy[off + i] = (i < n - 1) ? e[i] * e[i] : 0.0;
        }
    } else {
        split = -DBL_EPSILON;
        tryrac = 0;
    }
    double rtol1, rtol2;
    if (dsc.jobz != kV) { rtol1 = 4.0 * DBL_EPSILON; rtol2 = 4.0 * DBL_EPSILON; }
    else {
        rtol1 = fmin(sqrt(DBL_EPSILON), 1e-2 * kMinRelgap);
        rtol2 = fmax(fmin(rtol1 * 5.0E-3, kMinRelgap * 5.0E-6), 4.0 * DBL_EPSILON);
    }
    const double bsrtol = fmin(kMinRelgap * 1.0E-2, sqrt(DBL_EPSILON));
    const double rqtol = 2.0 * DBL_EPSILON;

    int irange = dsc.range;
    int il = dsc.il, iu = dsc.iu;
    if (irange == kInd && il == 1 && iu == n) irange = kAll;

    if (n == 1) { /* mrrr_val.c:193-211 (only reached for jobz = C) */
        if (tid == 0) {
            int in = (irange == kAll) || (irange == kVal && d[0] > vl && d[0] <= vu) ||
                     (irange == kInd && il == 1 && iu == 1);
            S.m = S.mtot = S.nvals = in;
            S.done = 1;
        }
        return;
    }
    // ---- E2, Gershgorin disks, pivmin (mrrr_val.c:214-241)
    double lemax = 0.0, lgl = d[0], lgu = d[0];
    if (tid == 0) e[n - 1] = 0.0;
    __syncthreads();
    for (int i = tid; i < n; i += PREP_NT) {
        double eabs = fabs(e[i]);
        double eold = (i > 0) ? fabs(e[i - 1]) : 0.0;
        e2[i] = e[i] * e[i];
        lemax = fmax(lemax, eabs);
        double tmp1 = eabs + eold;
        g[2 * i] = d[i] - tmp1;
        g[2 * i + 1] = d[i] + tmp1;
        lgl = fmin(lgl, g[2 * i]);
        lgu = fmax(lgu, g[2 * i + 1]);
    }
    const double emax = wg_reduce_max<PREP_NT>(lemax, shd);
    const double glraw = wg_reduce_min<PREP_NT>(lgl, shd);
    const double guraw = -wg_reduce_min<PREP_NT>(-lgu, shd);
    const double pivmin = DBL_MIN * fmax(1.0, emax * emax);
    const double spdiam = guraw - glraw;
    const double tnorm = fmax(fabs(glraw), fabs(guraw));
    const double gl = glraw - 2 * tnorm * DBL_EPSILON * n - 4 * pivmin;
    const double gu = guraw + 2 * tnorm * DBL_EPSILON * n + 4 * pivmin;
    if (irange == kVal && vl < gl && vu >= gu) irange = kAll;

    // ---- dlarra (odrra.f:90-123): split points, compacted into isplit (1-based ends)
    int nsplit = 0;
    for (int base = 0; base < n - 1; base += PREP_NT) {
        int i = base + tid;
        int f = 0;
        if (i < n - 1) {
            double eabs = fabs(e[i]);
            if (split < 0.0) f = (eabs <= fabs(split) * spdiam);
            else f = (eabs <= split * sqrt(fabs(d[i])) * sqrt(fabs(d[i + 1])));
        }
        int tot;
        int pos = wg_exscan<PREP_NT>(f, shi, &tot);
        if (f) {
            e[i] = 0.0;
            e2[i] = 0.0;
            isplit[off + nsplit + pos] = i + 1;
        }
        nsplit += tot;
    }
    if (tid == 0) isplit[off + nsplit] = n;
    nsplit += 1;
    __syncthreads();

    // ---- counts (mrrr_val.c:253-277) and the range-I bracket (:293-319)
    if (irange == kAll) { il = 1; iu = n; }
    else if (irange == kVal) {
        if (tid < 2) {
            int c = count_T_ijob1(n, d, e2, tid == 0 ? vl : vu, pivmin);
            if (tid == 0) sh_c0 = c; else sh_c1 = c;
        }
        __syncthreads();
        il = sh_c0 + 1;
        iu = sh_c1;
    }
    const int nvals = iu - il + 1;
    if (dsc.jobz == kC) {
        if (tid == 0) {
            S.m = S.mtot = S.nvals = nvals;
            S.il = il; S.iu = iu;
            S.done = 1;
        }
        return;
    }
    double wl, wu, wlu = 0.0, wul = 0.0;
    int nwl = 0, nwu = 0;
    if (irange == kInd) {
        const int itmax = (int)((log(tnorm + pivmin) - log(pivmin)) / log(2.0) + 2);
        const double atoli = 2 * (2 * DBL_MIN + 2 * pivmin);
        __shared__ Ijob3 st3[2];
        __shared__ int cnt3[2][1 << kL3];
        const int nval[2] = {il - 1, iu};
        if (tid < 2) {
            // both start from [gl, gu] with counts -1, n + 1; the first point is the
            // interval's own end (gl for il - 1, gu for iu)
            st3[tid].a1 = gl; st3[tid].a2 = gu; st3[tid].c = tid == 0 ? gl : gu;
            st3[tid].n1 = -1; st3[tid].n2 = n + 1; st3[tid].it = 0; st3[tid].done = itmax <= 0;
        }
        ijob3_pair(n, d, e2, pivmin, atoli, rqtol, itmax, nval, st3, cnt3);
        wl = st3[0].a1; wlu = st3[0].a2; nwl = st3[0].n1;
        wu = st3[1].a2; wul = st3[1].a1; nwu = st3[1].n2;
        vl = wl;
        vu = wu;
    } else if (irange == kVal) {
        wl = vl;
        wu = vu;
    } else {
        wl = gl;
        wu = gu;
        vl = gl;
        vu = gu;
    }

    // ---- per-block bounds and wanted index ranges (mrrr_val.c:340-418)
    int mtot = 0, maxbl = 0;
    for (int base = 0; base < nsplit; base += PREP_NT) {
        int jb = base + tid;
        int mbl = 0;
        BlockRec br;
        if (jb < nsplit) {
            int ibgn = (jb == 0) ? 0 : isplit[off + jb - 1];
            int iend = isplit[off + jb] - 1;
            int nbl = iend - ibgn + 1;
            br.bbeg = ibgn; br.bend = iend; br.mat = k; br.sigma = 0.0;
            br.ilow = 1; br.iupp = 0;
            br.gl = br.gu = br.glraw = br.guraw = d[ibgn];
            br.bl = br.bu = d[ibgn];
            br.nl = 0;
            br.nr = nbl;
            br.spdiam = 0.0;
            if (nbl == 1) {
                int in = (irange == kAll) || (irange == kVal && d[ibgn] > wl && d[ibgn] <= wu) ||
                         (irange == kInd && (d[ibgn] + pivmin) > wl && (d[ibgn] - pivmin) <= wu);
                mbl = in;
                br.ilow = 1; br.iupp = in;
                e[iend] = 0.0;
            } else {
                double bgl, bgu;
                if (nsplit == 1) { bgl = glraw; bgu = guraw; }
                else {
                    bgl = d[ibgn]; bgu = d[ibgn];
                    for (int i = ibgn; i <= iend; ++i) { bgl = fmin(bgl, g[2 * i]); bgu = fmax(bgu, g[2 * i + 1]); }
                }
                br.glraw = bgl; br.guraw = bgu;
                br.gl = bgl - 2 * tnorm * DBL_EPSILON * nbl - 4 * pivmin;
                br.gu = bgu + 2 * tnorm * DBL_EPSILON * nbl + 4 * pivmin;
                coarse_bounds(br, nbl, pivmin, &br.bl, &br.bu);
                br.nl = 0;
                br.nr = nbl;
                int ilow, iupp;
                if (irange == kAll) { ilow = 1; iupp = nbl; }
                else {
                    ilow = count_T_ijob1(nbl, d + ibgn, e2 + ibgn, fmax(wl, br.gl), pivmin) + 1;
                    iupp = count_T_ijob1(nbl, d + ibgn, e2 + ibgn, fmin(wu, br.gu), pivmin);
                }
                mbl = iupp - ilow + 1;
                if (mbl < 0) mbl = 0;
                br.ilow = ilow; br.iupp = iupp;
                if (mbl == 0) e[iend] = 0.0;
                else maxbl = max(maxbl, nbl);
            }
        }
        int tot;
        int pos = wg_exscan<PREP_NT>(mbl, shi, &tot);
        if (jb < nsplit) {
            br.wbeg = mtot + pos;
            br.wend = mtot + pos + mbl - 1;
            blk[off + jb] = br;
        }
        mtot += tot;
    }
    __syncthreads();
    // ---- slot records: iblock, Windex; 1x1 blocks get their eigenvalue now
    if (nsplit <= 64) {
        for (int jb = 0; jb < nsplit; ++jb) {
            BlockRec br = blk[off + jb];
            int nbl = br.bend - br.bbeg + 1;
            for (int s = br.wbeg + tid; s <= br.wend; s += PREP_NT) {
                iblock[off + s] = jb + 1;
                Windex[off + s] = (nbl == 1) ? 1 : br.ilow + (s - br.wbeg);
                if (nbl == 1) {
                    W[off + s] = d[br.bbeg];
                    Wcoarse[off + s] = d[br.bbeg];
                    Werr[off + s] = 0.0;
                }
            }
        }
    } else {
        for (int jb = tid; jb < nsplit; jb += PREP_NT) {
            BlockRec br = blk[off + jb];
            int nbl = br.bend - br.bbeg + 1;
            for (int s = br.wbeg; s <= br.wend; ++s) {
                iblock[off + s] = jb + 1;
                Windex[off + s] = (nbl == 1) ? 1 : br.ilow + (s - br.wbeg);
                if (nbl == 1) {
                    W[off + s] = d[br.bbeg];
                    Wcoarse[off + s] = d[br.bbeg];
                    Werr[off + s] = 0.0;
                }
            }
        }
    }
    int maxbl_all = (int)(-wg_reduce_min<PREP_NT>(-(double)maxbl, shd));
    if (tid == 0) {
        S.scale = scale;
        S.pivmin = pivmin;
        S.spdiam = spdiam;
        S.tnorm = tnorm;
        S.gl = gl; S.gu = gu; S.glraw = glraw; S.guraw = guraw;
        S.wl = wl; S.wu = wu; S.wlu = wlu; S.wul = wul;
        S.vl = vl; S.vu = vu;
        S.rtol1 = rtol1; S.rtol2 = rtol2; S.bsrtol = bsrtol; S.rqtol = rqtol; S.split = split;
        S.il = il; S.iu = iu; S.nvals = nvals;
        S.mtot = mtot;
        S.m = mtot;
        S.nsplit = nsplit;
        S.nwl = nwl; S.nwu = nwu;
        S.tryrac = tryrac;
        S.irange = irange;
        S.maxbl = maxbl_all;
    }
}

// ---------------------------------------------------------------------------------
// k_block_bracket: the range-I bracket of each block whose wanted index range is not
// its whole spectrum.  mrrr_val hands such a block to dlarrd with range I and the
// block's own index range ilow..iupp (mrrr_val.c:421, find_eigval_irreducible), and
// dlarrd first brackets those indices by dlaebz IJOB = 3 on the block (odrrd.f:343-393,
// from the block's Gershgorin interval fudged by FUDGE·tnorm·eps·n + FUDGE·2·pivmin, to
// RELTOL = bsrtol), then bisects from the block loop's interval cut to that bracket
// (odrrd.f:455-480), with the counts at its ends.  Those are the points the reference's
// bisection visits, so the coarse eigenvalues are its bits.  One workgroup per matrix,
// its blocks one after the other (ijob3_pair spreads each step over the workgroup).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(PREP_NT) void k_block_bracket(const MatDesc* __restrict__ md,
                                                           MatState* __restrict__ ms,
                                                           const double* __restrict__ D,
                                                           const double* __restrict__ E2,
                                                           BlockRec* blk) {
    __shared__ Ijob3 st3[2];
    __shared__ int cnt3[2][1 << kL3];
    __shared__ int sh_bad;
    const int k = blockIdx.x, tid = threadIdx.x;
    MatState& S = ms[k];
    if (S.done || S.status != kOk || S.irange == kAll) return;
    const int off = md[k].off, nsplit = S.nsplit;
    const double pivmin = S.pivmin, bsrtol = S.bsrtol;
    const double atoli = 2.0 * 2.0 * kSafmin + 2.0 * 2.0 * pivmin;
    if (tid == 0) sh_bad = 0;
    for (int jb = 0; jb < nsplit; ++jb) {
        __syncthreads();
        BlockRec br = blk[off + jb];
        const int nbl = br.bend - br.bbeg + 1;
        const int mbl = br.wend - br.wbeg + 1;
        if (nbl == 1 || mbl <= 0 || (br.ilow == 1 && br.iupp == nbl)) continue;   // uniform
        const double* d = D + off + br.bbeg;
        const double* e2 = E2 + off + br.bbeg;
        // odrrd.f:330-341 on the block alone
        const double tnorm = fmax(fabs(br.glraw), fabs(br.guraw));
        const double gl = br.glraw - 2.0 * tnorm * kEps * nbl - 2.0 * 2.0 * pivmin;
        const double gu = br.guraw + 2.0 * tnorm * kEps * nbl + 2.0 * 2.0 * pivmin;
        const int itmax = (int)((log(tnorm + pivmin) - log(pivmin)) / log(2.0)) + 2;
        const int nval[2] = {br.ilow - 1, br.iupp};
        if (tid < 2) {
            st3[tid].a1 = gl; st3[tid].a2 = gu; st3[tid].c = tid == 0 ? gl : gu;
            st3[tid].n1 = -1; st3[tid].n2 = nbl + 1; st3[tid].it = 0; st3[tid].done = itmax <= 0;
        }
        ijob3_pair(nbl, d, e2, pivmin, atoli, bsrtol, itmax, nval, st3, cnt3);
        if (tid == 0) {
            const double wl = st3[0].a1, wu = st3[1].a2;
            const int nwl = st3[0].n1, nwu = st3[1].n2;
            double bl = br.bl, bu = br.bu;   // the block loop's interval (odrrd.f:455-462)
            bool bad = nwl < 0 || nwl >= nbl || nwu < 1 || nwu > nbl;   // odrrd.f:384-390
            if (!bad) {
                if (bu < wl) bad = true;      // no wanted eigenvalue here after all
                bl = fmax(bl, wl);
                bu = fmin(bu, wu);
                if (bl >= bu) bad = true;
            }
            if (bad) {
                // the reference's dlarrd would return an error or fewer values than
                // mrrr_val counted: the problem fails (info 2)
                sh_bad = 1;
            } else {
                br.bl = bl;
                br.bu = bu;
                br.nl = count_T_ijob1(nbl, d, e2, bl, pivmin);   // IJOB = 1 at both ends
                br.nr = count_T_ijob1(nbl, d, e2, bu, pivmin);
                blk[off + jb] = br;
            }
        }
    }
    __syncthreads();
    if (tid == 0 && sh_bad) { S.status = kNoRoot; S.done = 1; }
}

// ---------------------------------------------------------------------------------
// k_coarse: eigenvalue approximations by bisection on T, one lane per wanted
// eigenvalue (dlarrd/dlaebz, odrrd.f:415-574, to RELTOL = bsrtol).
//
// Every eigenvalue of a block starts from the same interval, and a bisection step's
// midpoint depends only on the interval, so the first steps of all eigenvalues of a
// block visit the nodes of one binary tree: node t (root 1, children 2t for "the
// eigenvalue lies left of the midpoint", 2t + 1 otherwise).  k_coarse_tree counts the
// eigenvalues below the midpoint of every node of the top coarse_tree_depth levels
// once per block; k_coarse then takes those counts from the table and sweeps only
// below it.  Each lane's sequence of intervals, counts and tests is the one of its own
// bisection, so the results are the same bit for bit.
// ---------------------------------------------------------------------------------
//
// The table's depth is log2(nbl) - 1 when the whole block is wanted: each lane saves
// that many sweeps for about nbl / 2 counts taken once.  When only mbl eigenvalues of
// the block are wanted (range I or V), the table stops at log2(mbl) - 1 levels, or at 8
// (one count per lane of the workgroup) if that is more: its counts run one after the
// other on the workgroup, so levels that few lanes use would lengthen the critical path.
__device__ __forceinline__ int coarse_tree_depth(int nbl, int mbl) {
    if (nbl < 128) return 0;
    const int full = (31 - __clz(nbl)) - 1;
    const int want = mbl > 1 ? (31 - __clz(mbl)) - 1 : 0;
    return min(full, max(8, want));
}

// Initial interval of a block's coarse bisection: dlarrd called on the block alone
// (mrrr_val.c:1291-1295) with its own fudged bounds (odrrd.f:515-519, FUDGE = 2).
// Counts at the tree nodes 1 .. 2^depth - 1 of every block of at least 128 rows; one
// workgroup per block, tcount[off + bbeg + t - 1] = count at node t.
__device__ __forceinline__ void k_coarse_tree_one(int bx, const MatDesc* __restrict__ md,
                                                     const MatState* __restrict__ ms,
                                                     const double* __restrict__ D,
                                                     const double* __restrict__ E2,
                                                     const BlockRec* __restrict__ blk,
                                                     const int* __restrict__ smat,
                                                     const int* __restrict__ blist, int nblist,
                                                     int* tcount) {
    const int ib = bx;
    if (ib >= nblist) return;
    const int gb = blist[ib];
    const int k = smat[gb];
    const MatState& S = ms[k];
    const int off = md[k].off;
    if (S.done || gb - off >= S.nsplit) return;
    const BlockRec br = blk[gb];
    const int nbl = br.bend - br.bbeg + 1;
    if (br.wend < br.wbeg) return;
    const int depth = coarse_tree_depth(nbl, br.wend - br.wbeg + 1);
    if (depth == 0) return;
    const double pivmin = S.pivmin;
    const double* d = D + off + br.bbeg;
    const double* e2 = E2 + off + br.bbeg;
    double left0, right0;
    left0 = br.bl;
    right0 = br.bu;
    int sweeps = 0;
    for (int t = threadIdx.x + 1; t < (1 << depth); t += blockDim.x) {
        const int dt = 31 - __clz(t);
        double l = left0, r = right0;
        for (int b = dt - 1; b >= 0; --b) {
            const double mid = 0.5 * (l + r);
            if ((t >> b) & 1) l = mid;
            else r = mid;
        }
        const double mid = 0.5 * (l + r);
        tcount[off + br.bbeg + t - 1] = count_T_ijob2_w(nbl, d, e2, mid, pivmin);
        ++sweeps;
    }
    add_steps(kStepCoarse, (unsigned long long)sweeps * nbl);
}

__global__ __launch_bounds__(256) void k_coarse_tree(const MatDesc* __restrict__ md,
                                                     const MatState* __restrict__ ms,
                                                     const double* __restrict__ D,
                                                     const double* __restrict__ E2,
                                                     const BlockRec* __restrict__ blk,
                                                     const int* __restrict__ smat,
                                                     const int* __restrict__ blist, const int* __restrict__ nblist_p,
                                                     int* tcount) {
    const int nblist = *nblist_p;
    const int nblk_ = nblist;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_coarse_tree_one(bx, md, ms, D, E2, blk, smat, blist, nblist, tcount);
    }
}

__global__ __launch_bounds__(256) void k_coarse(const MatDesc* __restrict__ md,
                                                const MatState* __restrict__ ms,
                                                const double* __restrict__ D,
                                                const double* __restrict__ E2,
                                                const BlockRec* __restrict__ blk,
                                                const int* __restrict__ smat,
                                                const int* __restrict__ iblock,
                                                const int* __restrict__ Windex,
                                                const int* __restrict__ tcount, double* W,
                                                double* Werr, double* Wcoarse, int ntot) {
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= ntot) return;
    const int k = smat[gidx];
    const MatState& S = ms[k];
    const int off = md[k].off;
    const int s = gidx - off;
    if (S.done || s >= S.mtot) return;
    const BlockRec br = blk[off + iblock[gidx] - 1];
    const int nbl = br.bend - br.bbeg + 1;
    if (nbl == 1) return;
    const double pivmin = S.pivmin;
    const double* d = D + off + br.bbeg;
    const double* e2 = E2 + off + br.bbeg;
    const int i = Windex[gidx];
    // the block's own bounds and tolerances (odrrd.f:337-341)
    double left = br.bl, right = br.bu;
    const double atoli = 2.0 * 2.0 * kSafmin + 2.0 * 2.0 * pivmin;
    const double rtoli = S.bsrtol;
    const int itmax = (int)((log(right - left + pivmin) - log(pivmin)) / log(2.0)) + 2;
    const int depth = coarse_tree_depth(nbl, br.wend - br.wbeg + 1);
    const int* tc = tcount + off + br.bbeg - 1;   // tc[t]: count at tree node t
    int nl = br.nl, nr = br.nr, sweeps = 0, t = 1;
    for (int it = 0; it < itmax; ++it) {
        double tmp1 = fabs(right - left), tmp2 = fmax(fabs(right), fabs(left));
        if (tmp1 < fmax(fmax(atoli, pivmin), rtoli * tmp2)) break;
        double mid = 0.5 * (left + right);
        int c;
        if (it < depth) c = tc[t];
        else {
            c = count_T_ijob2_w(nbl, d, e2, mid, pivmin);
            ++sweeps;
        }
        c = min(nr, max(nl, c));
        if (c >= i) { right = mid; nr = c; t = 2 * t; }
        else { left = mid; nl = c; t = 2 * t + 1; }
    }
    add_steps(kStepCoarse, (unsigned long long)sweeps * nbl);
    const double w = 0.5 * (left + right);
    W[gidx] = w;
    Wcoarse[gidx] = w;
    Werr[gidx] = 0.5 * fabs(left - right);
}

// ---------------------------------------------------------------------------------
// k_block_list: the (matrix offset + block index) of every block of the matrices that
// take part in a phase, matrices in the order given (largest first), blocks in order;
// *count receives the length.  One workgroup: each thread counts a run of matrices,
// a workgroup scan places the runs.  which = 0: blocks for the shared top of the
// coarse bisection (matrices not done with a block of 128 rows or more); 1: blocks of
// the root representations (matrices not done whose status is ok).
// ---------------------------------------------------------------------------------
constexpr int kListNT = 1024;
__global__ __launch_bounds__(kListNT) void k_block_list(const MatDesc* __restrict__ md,
                                                        const MatState* __restrict__ ms,
                                                        const int* __restrict__ order, int nmat,
                                                        int which, int* list, int* count) {
    __shared__ int part[kListNT];
    const int t = threadIdx.x;
    const int per = (nmat + kListNT - 1) / kListNT;
    const int lo = min(nmat, t * per), hi = min(nmat, lo + per);
    auto blocks = [&](int k) -> int {
        const MatState& S = ms[k];
        if (S.done) return 0;
        if (which == 0) return S.maxbl >= 128 ? S.nsplit : 0;
        return S.status == kOk ? S.nsplit : 0;
    };
    int mine = 0;
    for (int j = lo; j < hi; ++j) mine += blocks(order[j]);
    part[t] = mine;
    __syncthreads();
    for (int o = 1; o < kListNT; o <<= 1) {   // inclusive scan
        const int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int pos = part[t] - mine;
    for (int j = lo; j < hi; ++j) {
        const int k = order[j];
        const int nb = blocks(k);
        for (int b = 0; b < nb; ++b) list[pos++] = md[k].off + b;
    }
    if (t == kListNT - 1) *count = part[t];
}

// Fills n ints with zero, grid-stride (a kernel, not a memset, so that a captured
// graph replays it like any other node).
__global__ __launch_bounds__(256) void k_fill0(int* a, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        a[i] = 0;
}

// ---------------------------------------------------------------------------------
// k_discard: range I surplus removal (mrrr_val.c:520-626), one thread per matrix.
// Runs before the root representation is formed; the slots of each block are
// re-packed.
// ---------------------------------------------------------------------------------
__global__ void k_discard(const MatDesc* __restrict__ md, MatState* ms, BlockRec* blk,
                          double* W, double* Werr, double* Wcoarse, int* iblock, int* Windex,
                          int nmat) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nmat) return;
    MatState& S = ms[k];
    if (S.done || S.irange != kInd || S.mtot <= S.nvals) return;
    const int off = md[k].off;
    double* w = W + off; double* we = Werr + off; double* wc = Wcoarse + off;
    int* ib = iblock + off; int* wi = Windex + off;
    int mp = S.mtot;
    int idiscl = S.il - 1 - S.nwl, idiscu = S.nwu - S.iu;
    if (idiscl > 0) {
        int im = 0;
        for (int je = 0; je < mp; ++je) {
            if (wc[je] < S.wlu && idiscl > 0) --idiscl;
            else { w[im] = w[je]; wc[im] = wc[je]; we[im] = we[je]; wi[im] = wi[je]; ib[im] = ib[je]; ++im; }
        }
        mp = im;
    }
    if (idiscu > 0) {
        int im = mp;
        for (int je = mp - 1; je >= 0; --je) {
            if (wc[je] > S.wul && idiscu > 0) --idiscu;
            else { --im; w[im] = w[je]; wc[im] = wc[je]; we[im] = we[je]; wi[im] = wi[je]; ib[im] = ib[je]; }
        }
        int jee = 0;
        for (int je = im; je < mp; ++je) {
            w[jee] = w[je]; wc[jee] = wc[je]; we[jee] = we[je]; wi[jee] = wi[je]; ib[jee] = ib[je]; ++jee;
        }
        mp -= im;
    }
    if (idiscl > 0 || idiscu > 0) {
        if (idiscl > 0) {
            double wkill = S.wu;
            for (int jd = 0; jd < idiscl; ++jd) {
                int iw = -1;
                for (int je = 0; je < mp; ++je)
                    if (ib[je] != 0 && (wc[je] < wkill || iw == -1)) { iw = je; wkill = wc[je]; }
                ib[iw] = 0;
            }
        }
        if (idiscu > 0) {
            double wkill = S.wl;
            for (int jd = 0; jd < idiscu; ++jd) {
                int iw = -1;
                for (int je = 0; je < mp; ++je)
                    if (ib[je] != 0 && (wc[je] > wkill || iw == -1)) { iw = je; wkill = wc[je]; }
                ib[iw] = 0;
            }
        }
        int im = 0;
        for (int je = 0; je < mp; ++je)
            if (ib[je] != 0) { w[im] = w[je]; wc[im] = wc[je]; we[im] = we[je]; wi[im] = wi[je]; ib[im] = ib[je]; ++im; }
        mp = im;
    }
    S.mtot = mp;
    S.m = mp;
    // re-derive the slot range of every block from the packed iblock list
    for (int jb = 0; jb < S.nsplit; ++jb) { blk[off + jb].wbeg = 0; blk[off + jb].wend = -1; }
    for (int s = 0; s < mp; ++s) {
        BlockRec& br = blk[off + ib[s] - 1];
        if (br.wend < br.wbeg) { br.wbeg = s; br.wend = s; br.ilow = wi[s]; }
        else br.wend = s;
        br.iupp = wi[s];
    }
    if (mp < S.nvals) S.status = kTooFew;
}

// ---------------------------------------------------------------------------------
// k_root: gaps from the coarse values, then the root representation of a block
// (compute_root_rrr, mrrr_val.c:717-948) and the set-up of the refinement
// (mrrr_val.c:1003-1011).  One 64-lane workgroup per block slot.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void k_root_one(int bx, const MatDesc* __restrict__ md, MatState* ms,
                                             BlockRec* blk, const int* __restrict__ smat,
                                             const int* __restrict__ blist, double* D, double* E,
                                             double* DLr, double* DLLr, double* W, double* Werr,
                                             double* Wgap, int nblist) {
    if ((int)bx >= nblist) return;
    const int gb = blist[bx];
    const int k = smat[gb];
    MatState& S = ms[k];
    const int off = md[k].off;
    const int jb = gb - off;
    if (S.done || jb >= S.nsplit) return;
    BlockRec& br = blk[gb];
    const int nbl = br.bend - br.bbeg + 1;
    const int wbgn = off + br.wbeg, wend = off + br.wend;
    const int mbl = br.wend - br.wbeg + 1;
    if (nbl == 1 || mbl <= 0) return;
    const int lane = threadIdx.x;
    double* d = D + off + br.bbeg;
    double* e = E + off + br.bbeg;
    double* dl = DLr + off + br.bbeg;
    double* dll = DLLr + off + br.bbeg;
    const double pivmin = S.pivmin;
    __shared__ double sh_sigma, sh_spdiam, sh_tau, sh_dmax;
    __shared__ int sh_fail, sh_sgndef;
    // rows of the trial factorisation staged through LDS (d, e in; D, L out), so that the
    // serial loop neither waits on global loads nor reads back its own global stores
    constexpr int RC = 512;
    __shared__ double sd[RC], se[RC], sdp[RC], sll[RC];
    // gaps of the coarse values (mrrr_val.c:431-435)
    for (int s = wbgn + lane; s <= wend; s += 64) {
        if (s < wend) Wgap[s] = fmax(0.0, (W[s + 1] - Werr[s + 1]) - (W[s] + Werr[s]));
        else Wgap[s] = fmax(0.0, S.wu - (W[s] + Werr[s]));
    }
    __syncthreads();
    const double isleft = fmax(br.gl, (W[wbgn] - Werr[wbgn]) - 100.0 * kEps * fabs(W[wbgn] - Werr[wbgn]));
    const double isright = fmin(br.gu, (W[wend] + Werr[wend]) + 100.0 * kEps * fabs(W[wend] - Werr[wend]));
    int lcnt = 0, rcnt = 0;
    if (mbl > 1) {
        // dlarrc 'T' (odrrc.f:95-113) at the two quarter points, rows staged through LDS
        const double tmp0 = fmin(isright, S.vu) - fmax(isleft, S.vl);
        const double s1 = fmax(isleft, S.vl) + 0.25 * tmp0;
        const double s2 = fmin(isright, S.vu) - 0.25 * tmp0;
        double lp = d[0] - s1, rp = d[0] - s2;
        lcnt = (lp <= 0.0); rcnt = (rp <= 0.0);
        for (int base = 0; base < nbl - 1; base += RC) {
            const int cnt = min(RC, nbl - 1 - base);
            __syncthreads();
            for (int j = lane; j < cnt; j += 64) {
                sd[j] = d[base + j + 1];
                se[j] = e[base + j];
            }
            __syncthreads();
            if (lane == 0) {
                for (int j = 0; j < cnt; ++j) {
                    const double t = se[j] * se[j];
                    lp = (sd[j] - s1) - t / lp;
                    rp = (sd[j] - s2) - t / rp;
                    lcnt += (lp <= 0.0); rcnt += (rp <= 0.0);
                }
            }
        }
    }
    if (lane == 0) {
        const double gl = br.gl, gu = br.gu, vl = S.vl, vu = S.vu;
        const int ilow = br.ilow, iupp = br.iupp;
        double spdiam = gu - gl;
        double sigma, tau;
        int sgndef;
        if (mbl == 1) { sigma = gl; sgndef = 1; }
        else if ((lcnt - ilow) >= (iupp - rcnt)) { sigma = fmax(isleft, vl); sgndef = 1; }
        else { sigma = fmin(isright, vu); sgndef = -1; }
        if (mbl > 1) {
            double clwdth = W[wend] + Werr[wend] - W[wbgn] - Werr[wbgn];
            double avgap = fabs(clwdth / (double)(wend - wbgn));
            if (sgndef > 0) { tau = fmax(Wgap[wbgn], avgap); tau *= 0.5; tau = fmax(tau, Werr[wbgn]); }
            else { tau = fmax(Wgap[wend - 1], avgap); tau *= 0.5; tau = fmax(tau, Werr[wend]); }
        } else tau = Werr[wbgn];
        sh_sigma = sigma;
        sh_tau = tau;
        sh_sgndef = sgndef;
        sh_spdiam = spdiam;
    }
    __syncthreads();
    {
        // the tries at sigma (every lane holds the same sigma, tau and decisions)
        double sigma = sh_sigma, tau = sh_tau;
        const double spdiam = sh_spdiam;
        const int sgndef = sh_sgndef;
        int ok = 0;
        for (int jtry = 0; jtry < kMaxTryRoot; ++jtry) {
            // D of the trial factorisation in dl, L in dll
            double dpivot = d[0] - sigma;
            double dmax = fabs(dpivot);
            if (lane == 0) dl[0] = dpivot;
            for (int base = 0; base < nbl - 1; base += RC) {
                const int cnt = min(RC, nbl - 1 - base);
                for (int j = lane; j < cnt; j += 64) {
                    sd[j] = d[base + j + 1];
                    se[j] = e[base + j];
                }
                __syncthreads();
                if (lane == 0) {
                    for (int j = 0; j < cnt; ++j) {
                        const double invd = 1.0 / dpivot;
                        const double t = se[j] * invd;
                        sll[j] = t;
                        dpivot = (sd[j] - sigma) - t * se[j];
                        sdp[j] = dpivot;
                        dmax = fmax(dmax, fabs(dpivot));
                    }
                }
                __syncthreads();
                for (int j = lane; j < cnt; j += 64) {
                    dll[base + j] = sll[j];
                    dl[base + j + 1] = sdp[j];
                }
            }
            if (lane == 0) sh_dmax = dmax;
            __syncthreads();
            dmax = sh_dmax;
            if (dmax > kMaxGrowth * spdiam) {
                if (jtry == kMaxTryRoot - 2) {
                    if (sgndef == 1) sigma = br.gl - kFudgeRoot * spdiam * kEps * nbl - kFudgeRoot * 2.0 * pivmin;
                    else sigma = br.gu + kFudgeRoot * spdiam * kEps * nbl + kFudgeRoot * 2.0 * pivmin;
                } else if (jtry == kMaxTryRoot - 1) {
                    break;
                } else {
                    sigma -= sgndef * tau;
                    tau *= 2.0;
                }
            } else { ok = 1; break; }
        }
        if (lane == 0) {
            sh_fail = !ok;
            sh_sigma = sigma;
        }
    }
    __syncthreads();
    if (sh_fail) {
        if (lane == 0) S.status = kNoRoot;
        return;
    }
    const double sigma = sh_sigma;
    // D <- D_try, L <- L_try, perturbed by (1 + 8 eps u) with u from dlarnv(iseed=1,1,1,1)
    // (mrrr_val.c:928-941); entry i of the random vector is seed * a^(i+1).
    const unsigned long long seed0 = (1ULL << 36) | (1ULL << 24) | (1ULL << 12) | 1ULL;
    const unsigned long long step = lcg_pow(64);
    unsigned long long xd = (seed0 * lcg_pow(lane + 1)) & kMask48;
    unsigned long long xl = (seed0 * lcg_pow(nbl + lane + 1)) & kMask48;
    const bool perturb = (mbl > 1);
    for (int i = lane; i < nbl; i += 64) {
        double dv = dl[i];
        double lv = (i < nbl - 1) ? dll[i] : 0.0;
        if (perturb) {
            if (i < nbl - 1) {
                dv *= 1.0 + kEps * kRandFactor * lcg_uniform_pm1(xd);
                lv *= 1.0 + kEps * kRandFactor * lcg_uniform_pm1(xl);
            } else {
                unsigned long long xlast = (seed0 * lcg_pow(2 * nbl)) & kMask48;
                dv *= 1.0 + kEps * kRandFactor * lcg_uniform_pm1(xlast);
            }
        }
        xd = (xd * step) & kMask48;
        xl = (xl * step) & kMask48;
        d[i] = dv;
        if (i < nbl - 1) e[i] = lv;
    }
    __syncthreads();
    if (lane == 0) {
        e[nbl - 1] = sigma;
        br.sigma = sigma;
        br.spdiam = sh_spdiam;
    }
    for (int i = lane; i < nbl; i += 64) {
        if (i < nbl - 1) {
            double t = d[i] * e[i];
            dl[i] = t;
            dll[i] = t * e[i];
        } else {
            dl[i] = 0.0;
            dll[i] = 0.0;
        }
    }
    for (int s = wbgn + lane; s <= wend; s += 64) {
        W[s] -= sigma;
        Werr[s] += fabs(W[s]) * DBL_EPSILON;
    }
}

__global__ __launch_bounds__(64) void k_root(const MatDesc* __restrict__ md, MatState* ms,
                                             BlockRec* blk, const int* __restrict__ smat,
                                             const int* __restrict__ blist, double* D, double* E,
                                             double* DLr, double* DLLr, double* W, double* Werr,
                                             double* Wgap, const int* __restrict__ nblist_p) {
    const int nblist = *nblist_p;
    const int nblk_ = nblist;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_root_one(bx, md, ms, blk, smat, blist, D, E, DLr, DLLr, W, Werr, Wgap, nblist);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// k_refine_root: dlarrb on the root representation, one lane per eigenvalue
// (refine_eigval_approx, mrrr_val.c:1075-1084; the whole block is one call, so its
// first eigenvalue takes its own right gap as left gap, odrrb.f:163-170).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_refine_root(const MatDesc* __restrict__ md,
                                                     const MatState* __restrict__ ms,
                                                     const BlockRec* __restrict__ blk,
                                                     const int* __restrict__ smat,
                                                     const int* __restrict__ iblock,
                                                     const int* __restrict__ Windex,
                                                     const double* __restrict__ D,
                                                     const double* __restrict__ DLLr, double* W,
                                                     double* Werr, const double* __restrict__ Wgap,
                                                     double* Wtmp, double* Werrtmp, int ntot) {
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= ntot) return;
    const int k = smat[gidx];
    const MatState& S = ms[k];
    const int off = md[k].off;
    const int s = gidx - off;
    if (S.done || S.status != kOk || s >= S.mtot) return;
    const BlockRec br = blk[off + iblock[gidx] - 1];
    const int nbl = br.bend - br.bbeg + 1;
    double w = W[gidx], we = Werr[gidx];
    if (nbl > 1) {
        double rgap = Wgap[gidx];
        double lgap = (s == br.wbeg) ? rgap : Wgap[gidx - 1];
        double gap = fmin(lgap, rgap);
        int sw = dlarrb1_w(nbl, D + off + br.bbeg, DLLr + off + br.bbeg, Windex[gidx], S.rtol1,
                           S.rtol2, gap, &w, &we, S.pivmin, br.spdiam, nbl);
        add_steps(kStepRefine, (unsigned long long)sw * nbl);
    }
    Wtmp[gidx] = w;
    Werrtmp[gidx] = we;
}

// Gaps after the refinement (odrrb.f:286-291, mrrr_val.c:1084-1086).
__global__ __launch_bounds__(256) void k_gaps_root(const MatDesc* __restrict__ md,
                                                   const MatState* __restrict__ ms,
                                                   const BlockRec* __restrict__ blk,
                                                   const int* __restrict__ smat,
                                                   const int* __restrict__ iblock,
                                                   const double* __restrict__ Wtmp,
                                                   const double* __restrict__ Werrtmp, double* W,
                                                   double* Werr, double* Wgap, int ntot) {
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= ntot) return;
    const int k = smat[gidx];
    const MatState& S = ms[k];
    const int off = md[k].off;
    const int s = gidx - off;
    if (S.done || S.status != kOk || s >= S.mtot) return;
    const BlockRec br = blk[off + iblock[gidx] - 1];
    const int nbl = br.bend - br.bbeg + 1;
    W[gidx] = Wtmp[gidx];
    Werr[gidx] = Werrtmp[gidx];
    if (nbl == 1) return;
    double w = Wtmp[gidx], we = Werrtmp[gidx];
    if (s < br.wend) Wgap[gidx] = fmax(0.0, Wtmp[gidx + 1] - Werrtmp[gidx + 1] - w - we);
    else Wgap[gidx] = fmax(0.0, (S.vu - br.sigma) - (w + we));
}

// jobz = N: eigenvalues of the shifted root representations back to T (mrrr.c:263-269).
__global__ __launch_bounds__(256) void k_shift_n(const MatDesc* __restrict__ md,
                                                 const MatState* __restrict__ ms,
                                                 const BlockRec* __restrict__ blk,
                                                 const int* __restrict__ smat,
                                                 const int* __restrict__ iblock, double* W,
                                                 int ntot) {
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= ntot) return;
    const int k = smat[gidx];
    const MatState& S = ms[k];
    const int off = md[k].off;
    if (S.done || S.status != kOk || md[k].jobz != kN || gidx - off >= S.m) return;
    const BlockRec br = blk[off + iblock[gidx] - 1];
    W[gidx] += br.sigma;
}

// dlarrj refinement to high relative accuracy on the original T (mrrr.c:452-520),
// one lane per eigenvalue.
__global__ __launch_bounds__(256) void k_rrj(const MatDesc* __restrict__ md,
                                             const MatState* __restrict__ ms,
                                             const BlockRec* __restrict__ blk,
                                             const int* __restrict__ smat,
                                             const int* __restrict__ iblock,
                                             const int* __restrict__ Windex,
                                             const double* __restrict__ Dcopy,
                                             const double* __restrict__ E2copy, double* W,
                                             double* Werr, int ntot) {
    const int gidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (gidx >= ntot) return;
    const int k = smat[gidx];
    const MatState& S = ms[k];
    const int off = md[k].off;
    if (S.done || S.status != kOk || !S.tryrac || gidx - off >= S.m) return;
    const BlockRec br = blk[off + iblock[gidx] - 1];
    const int nbl = br.bend - br.bbeg + 1;
    double w = W[gidx], we = Werr[gidx];
    int sw = dlarrj1_w(nbl, Dcopy + off + br.bbeg, E2copy + off + br.bbeg, Windex[gidx],
                       4 * DBL_EPSILON, &w, &we, S.pivmin, S.spdiam);
    add_steps(kStepRrj, (unsigned long long)sw * nbl);
    W[gidx] = w;
    Werr[gidx] = we;
}

// ---------------------------------------------------------------------------------
// k_zindex: column of every eigenvector (init_zindex, mrrr_vec.c:380-445): order by
// lambda + sigma across blocks, by local index within a block.  One workgroup per
// matrix; a rank is formed by binary search in every other block.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ bool zkey_before(double la, int ia, int ba, double lb, int ib, int bb) {
    // true if (lb, ib) of block bb sorts before (la, ia) of block ba (bb != ba).  The
    // reference's comparator (mrrr_vec.c:433-442) orders by lambda, then by local
    // index, and leaves exact ties of both unordered; the block index settles those
    // here so that the columns form a permutation.
    if (lb < la) return true;
    if (lb > la) return false;
    if (ib != ia) return ib < ia;
    return bb < ba;
}

__global__ __launch_bounds__(256) void k_zindex(const MatDesc* __restrict__ md,
                                                MatState* __restrict__ ms,
                                                const BlockRec* __restrict__ blk,
                                                const double* __restrict__ W,
                                                const int* __restrict__ iblock,
                                                const int* __restrict__ Windex, int* Zidx) {
    const int k = blockIdx.x;
    MatState& S = ms[k];
    if (S.done || S.status != kOk || md[k].jobz != kV) return;
    const int off = md[k].off, m = S.m, nsplit = S.nsplit;
    if (m > md[k].mcols) {
        // more eigenvectors than the caller's Z holds: the problem fails (info 2) here,
        // before the tree walk gives it any task, so no column of Z or Zsupp is written
        __syncthreads();   // every thread has read the status before it changes
        if (threadIdx.x == 0) { S.status = kTooFew; S.done = 1; }
        return;
    }
    for (int s = threadIdx.x; s < m; s += blockDim.x) {
        if (nsplit == 1) { Zidx[off + s] = s; continue; }
        const int b = iblock[off + s] - 1;
        const BlockRec mb = blk[off + b];
        const double lam = W[off + s] + mb.sigma;
        const int il = Windex[off + s];
        int rank = s - mb.wbeg;   // entries of the own block before it
        for (int ob = 0; ob < nsplit; ++ob) {
            if (ob == b) continue;
            const BlockRec o = blk[off + ob];
            if (o.wend < o.wbeg) continue;
            int lo = o.wbeg, hi = o.wend + 1;   // first entry not before (lam, il)
            while (lo < hi) {
                int mid = (lo + hi) >> 1;
                if (zkey_before(lam, il, b, W[off + mid] + o.sigma, Windex[off + mid], ob)) lo = mid + 1;
                else hi = mid;
            }
            rank += lo - o.wbeg;
        }
        Zidx[off + s] = rank;
    }
}

// ---------------------------------------------------------------------------------
// Representation tree.  Level 0 = the root of each block (process_cluster.c:265-281
// refine_eigvals at depth 0, then PMR_create_subtasks); deeper clusters carry their
// own representation in a per-level arena.
// ---------------------------------------------------------------------------------
struct TreeOut {
    STask* S;               // indexed by global eigenvalue slot space: block b's
                            // singletons of a level fill [wseg_b, wseg_b + bcnt[b])
    ClTask* C;
    int* eig_slot;          // next level: slot of each eigen entry
    int* eig_cl;            // next level: cluster of each eigen entry
    int* bcnt;              // singletons emitted per block in this level
    LevelCounters* cnt;
    const int* halt;        // nonzero: the representation pool overflowed, emit nothing
    int capS, capC, capEig;
};

// Child cluster record of a level (PMR_create_subtasks, process_cluster.c:560-600).
__device__ __forceinline__ ClTask make_cluster(int f, int l, int cb, double cl_lgap, int depth,
                                               int mat, int gblk, int wseg, int bbeg, int bsize,
                                               double bl_spdiam, const double* rD, const double* rL,
                                               const double* rDL, const double* rDLL,
                                               const double* __restrict__ Wgap, int ebase) {
    ClTask c;
    c.cb = f; c.ce = l; c.mat = mat; c.depth = depth + 1;
    c.bbeg = bbeg; c.bsize = bsize; c.bl_spdiam = bl_spdiam;
    c.lgap = (f == cb) ? cl_lgap : Wgap[f - 1];
    c.pD = rD; c.pL = rL; c.pDL = rDL; c.pDLL = rDLL;
    c.cD = c.cL = c.cDL = c.cDLL = nullptr;
    c.tau = 0.0;
    c.shift = 0;
    c.arena_off = -1;
    c.eig_base = ebase;
    c.gblk = gblk;
    c.wseg = wseg;
    return c;
}

// Walk one cluster's eigenvalues and emit singleton and child-cluster tasks
// (PMR_create_subtasks, process_cluster.c:438-629).  Every singleton is its own
// task, the schedule the reference takes when max_size = 1 (:481-482).
__device__ void classify_cluster(int cb, int ce, int depth, int mat, int gblk, int wseg, int bbeg,
                                 int bsize, double bl_spdiam, double cl_lgap, const double* rD,
                                 const double* rL, const double* rDL, const double* rDLL,
                                 const double* __restrict__ Wgap, const double* __restrict__ Wsh,
                                 TreeOut out) {
    // pass 1: count
    int nS = 0, nC = 0, nE = 0;
    {
        int new_first = cb;
        for (int i = cb; i <= ce; ++i) {
            if (!(i == ce || Wgap[i] >= kMinRelgap * fabs(Wsh[i]))) continue;
            int new_last = i;
            int new_size = new_last - new_first + 1;
            if (new_size == 1) ++nS;
            else if (depth == 0 && new_size > 3) {
                int cl_first = new_first;
                double avggap = bl_spdiam / (bsize - 1);
                for (int kk = new_first + 1; kk < new_last; ++kk) {
                    int cl_last;
                    if (kk == new_last - 1) cl_last = new_last;
                    else if (kk != cl_first && Wgap[kk] > 2.0 * avggap) cl_last = kk;
                    else continue;
                    ++nC; nE += cl_last - cl_first + 1;
                    cl_first = kk + 1;
                }
            } else { ++nC; nE += new_size; }
            new_first = i + 1;
        }
    }
    if (nS) atomicAdd(&out.cnt->nS, nS);
    int sbase = nS ? wseg + atomicAdd(&out.bcnt[gblk], nS) : 0;
    int cbase = nC ? atomicAdd(&out.cnt->nC, nC) : 0;
    int ebase = nE ? atomicAdd(&out.cnt->nEig, nE) : 0;
    if (sbase + nS > out.capS || cbase + nC > out.capC || ebase + nE > out.capEig) {
        atomicOr(&out.cnt->err, 1);
        return;
    }
    if (nC) {
        atomicAdd((unsigned long long*)&out.cnt->arena, (unsigned long long)(4LL * bsize * nC));
        atomicMax(&out.cnt->maxbl, bsize);
    }
    if (nS) atomicMax(&out.cnt->maxbl, bsize);
    // pass 2: emit
    int new_first = cb;
    for (int i = cb; i <= ce; ++i) {
        if (!(i == ce || Wgap[i] >= kMinRelgap * fabs(Wsh[i]))) continue;
        int new_last = i;
        int new_size = new_last - new_first + 1;
        if (new_size == 1) {
            STask t;
            t.slot = new_first; t.mat = mat; t.bbeg = bbeg; t.bsize = bsize;
            t.gblk = gblk; t.wseg = wseg;
            t.bl_spdiam = bl_spdiam;
            t.lgap = (new_first == cb) ? cl_lgap : Wgap[new_first - 1];
            t.D = rD; t.L = rL; t.DL = rDL; t.DLL = rDLL;
            out.S[sbase++] = t;
        } else {
            auto emit = [&](int f, int l) {
                out.C[cbase] = make_cluster(f, l, cb, cl_lgap, depth, mat, gblk, wseg, bbeg, bsize,
                                            bl_spdiam, rD, rL, rDL, rDLL, Wgap, ebase);
                for (int j = f; j <= l; ++j) {
                    out.eig_slot[ebase] = j;
                    out.eig_cl[ebase] = cbase;
                    ++ebase;
                }
                ++cbase;
            };
            if (depth == 0 && new_size > 3) {
                int cl_first = new_first;
                double avggap = bl_spdiam / (bsize - 1);
                for (int kk = new_first + 1; kk < new_last; ++kk) {
                    int cl_last;
                    if (kk == new_last - 1) cl_last = new_last;
                    else if (kk != cl_first && Wgap[kk] > 2.0 * avggap) cl_last = kk;
                    else continue;
                    emit(cl_first, cl_last);
                    cl_first = kk + 1;
                }
            } else emit(new_first, new_last);
        }
        new_first = i + 1;
    }
}

// Level 0: one wave per block.  The eigenvalues of the block are walked in tiles of
// 64; a lane whose index ends a group (the gap test of process_cluster.c:500-503)
// takes that group: a singleton, a cluster, or at the root the sub-clusters cut at
// gaps above twice the average gap (process_cluster.c:520-556).  Counts are reduced
// over the wave and positions come from wave scans, so the tasks of a block are
// written in index order, as the sequential walk writes them.  1x1 blocks become
// singletons of size 1 (mrrr_vec.c:330-346).
__device__ __forceinline__ int wave_exscan64(int v, int* total) {
    const int lane = threadIdx.x & 63;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int t = __shfl_up(x, o, 64);
        if (lane >= o) x += t;
    }
    *total = __shfl(x, 63, 64);
    return x - v;
}

__device__ __forceinline__ void k_classify_root_one(int bx, const MatDesc* __restrict__ md,
                                                      const MatState* __restrict__ ms,
                                                      const BlockRec* __restrict__ blk,
                                                      const int* __restrict__ smat,
                                                      const int* __restrict__ blist,
                                                      const double* __restrict__ gers,
                                                      const double* __restrict__ D,
                                                      const double* __restrict__ E,
                                                      const double* __restrict__ DLr,
                                                      const double* __restrict__ DLLr, double* W,
                                                      const double* __restrict__ Werr,
                                                      const double* __restrict__ Wgap, double* Wsh,
                                                      TreeOut out, int nblist) {
    __shared__ int sh_base[3];
    const int ib = bx;
    const int lane = threadIdx.x;
    if (ib >= nblist) return;
    const int gb = blist[ib];
    const int k = smat[gb];
    const MatState& S = ms[k];
    const MatDesc& dsc = md[k];
    const int off = dsc.off;
    const int jb = gb - off;
    if (S.done || S.status != kOk || dsc.jobz != kV || jb >= S.nsplit) return;
    const BlockRec br = blk[gb];
    const int nbl = br.bend - br.bbeg + 1;
    const int mbl = br.wend - br.wbeg + 1;
    if (mbl <= 0) return;
    const int cb = off + br.wbeg, ce = off + br.wend;
    const double* rD = D + off + br.bbeg;
    const double* rL = E + off + br.bbeg;
    const double* rDL = DLr + off + br.bbeg;
    const double* rDLL = DLLr + off + br.bbeg;
    const int wseg = off + br.wbeg;
    if (nbl == 1) {
        if (lane != 0) return;
        atomicAdd(&out.cnt->nS, 1);
        int pos = wseg + atomicAdd(&out.bcnt[gb], 1);
        atomicMax(&out.cnt->maxbl, 1);
        if (pos >= out.capS) { atomicOr(&out.cnt->err, 1); return; }
        STask t;
        t.slot = cb; t.mat = k; t.bbeg = br.bbeg; t.bsize = 1; t.bl_spdiam = 0.0; t.lgap = 0.0;
        t.gblk = gb; t.wseg = wseg;
        t.D = rD; t.L = rL; t.DL = rDL; t.DLL = rDLL;
        out.S[pos] = t;
        Wsh[cb] = W[cb];
        return;
    }
    // refine_eigvals at depth 0 (process_cluster.c:265-281): the spread of the block
    // from gersch rows bl_begin .. bl_end-1, Wshifted = W, W += sigma.
    const double* g = gers + 2 * (size_t)off;
    double gl = g[2 * br.bbeg], gu = g[2 * br.bbeg + 1];
    for (int i = br.bbeg + 1 + lane; i < br.bend; i += 64) {
        gl = fmin(gl, g[2 * i]);
        gu = fmax(gu, g[2 * i + 1]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        gl = fmin(gl, __shfl_xor(gl, o, 64));
        gu = fmax(gu, __shfl_xor(gu, o, 64));
    }
    const double bl_spdiam = gu - gl;
    const double sigma = br.sigma;
    // lgap of the root task (mrrr_vec.c:357), from the unshifted value of the first slot
    const double lgap = fmax(0.0, (W[cb] + sigma) - Werr[cb] - S.vl);
    const double avggap = bl_spdiam / (nbl - 1);
    // group end test; the shifted value of slot i is the incoming W[i]
    auto ends = [&](int i, double wi) { return i == ce || Wgap[i] >= kMinRelgap * fabs(wi); };
    // the sub-cluster cuts of a root group [f, l] of more than 3 (process_cluster.c:520-556)
    auto cut = [&](int kk, int cl_first, int l) {
        return kk == l - 1 || (kk != cl_first && Wgap[kk] > 2.0 * avggap);
    };
    // pass 1: counts, and W / Wsh updated in place (each slot by the lane that reads it)
    int nS = 0, nC = 0, nE = 0;
    int first = cb;
    for (int t0 = cb; t0 <= ce; t0 += 64) {
        const int i = t0 + lane;
        bool e = false;
        if (i <= ce) {
            const double wi = W[i];
            e = ends(i, wi);
            Wsh[i] = wi;
            W[i] = wi + sigma;
        }
        const unsigned long long m = __ballot(e);
        if (e) {
            const unsigned long long below = m & ((1ull << lane) - 1ull);
            const int f = below ? t0 + 64 - __clzll(below) : first;
            const int size = i - f + 1;
            if (size == 1) ++nS;
            else if (size > 3) {
                int cl_first = f;
                for (int kk = f + 1; kk < i; ++kk) {
                    if (!cut(kk, cl_first, i)) continue;
                    const int cl_last = (kk == i - 1) ? i : kk;
                    ++nC; nE += cl_last - cl_first + 1;
                    cl_first = kk + 1;
                }
            } else { ++nC; nE += size; }
        }
        if (m) first = t0 + 64 - __clzll(m);
    }
    for (int o = 32; o > 0; o >>= 1) {
        nS += __shfl_xor(nS, o, 64);
        nC += __shfl_xor(nC, o, 64);
        nE += __shfl_xor(nE, o, 64);
    }
    const int tS = nS, tC = nC, tE = nE;
    if (lane == 0) {
        int sb = 0, cbs = 0, eb = 0;
        if (tS) { atomicAdd(&out.cnt->nS, tS); sb = atomicAdd(&out.bcnt[gb], tS); }
        if (tC) cbs = atomicAdd(&out.cnt->nC, tC);
        if (tE) eb = atomicAdd(&out.cnt->nEig, tE);
        bool bad = wseg + sb + tS > out.capS || cbs + tC > out.capC || eb + tE > out.capEig;
        if (bad) atomicOr(&out.cnt->err, 1);
        if (tC) {
            atomicAdd((unsigned long long*)&out.cnt->arena, (unsigned long long)(4LL * nbl * tC));
            atomicMax(&out.cnt->maxbl, nbl);
        }
        if (tS) atomicMax(&out.cnt->maxbl, nbl);
        sh_base[0] = bad ? -1 : wseg + sb;
        sh_base[1] = cbs;
        sh_base[2] = eb;
    }
    __syncthreads();
    if (sh_base[0] < 0) return;
    int sbase = sh_base[0], cbase = sh_base[1], ebase = sh_base[2];
    // pass 2: emit in index order
    first = cb;
    for (int t0 = cb; t0 <= ce; t0 += 64) {
        const int i = t0 + lane;
        bool e = false;
        if (i <= ce) e = ends(i, Wsh[i]);
        const unsigned long long m = __ballot(e);
        int f = first, cS = 0, cC = 0, cE = 0;
        if (e) {
            const unsigned long long below = m & ((1ull << lane) - 1ull);
            if (below) f = t0 + 64 - __clzll(below);
            const int size = i - f + 1;
            if (size == 1) cS = 1;
            else if (size > 3) {
                int cl_first = f;
                for (int kk = f + 1; kk < i; ++kk) {
                    if (!cut(kk, cl_first, i)) continue;
                    const int cl_last = (kk == i - 1) ? i : kk;
                    ++cC; cE += cl_last - cl_first + 1;
                    cl_first = kk + 1;
                }
            } else { cC = 1; cE = size; }
        }
        int aS, aC, aE;
        int oS = sbase + wave_exscan64(cS, &aS);
        int oC = cbase + wave_exscan64(cC, &aC);
        int oE = ebase + wave_exscan64(cE, &aE);
        if (cS) {
            STask t;
            t.slot = f; t.mat = k; t.bbeg = br.bbeg; t.bsize = nbl;
            t.gblk = gb; t.wseg = wseg;
            t.bl_spdiam = bl_spdiam;
            t.lgap = (f == cb) ? lgap : Wgap[f - 1];
            t.D = rD; t.L = rL; t.DL = rDL; t.DLL = rDLL;
            out.S[oS] = t;
        }
        if (cC) {
            auto emit = [&](int cf, int cl) {
                for (int j = cf; j <= cl; ++j) {
                    out.eig_slot[oE] = j;
                    out.eig_cl[oE] = oC;
                    ++oE;
                }
                out.C[oC] = make_cluster(cf, cl, cb, lgap, 0, k, gb, wseg, br.bbeg, nbl, bl_spdiam,
                                         rD, rL, rDL, rDLL, Wgap, oE - (cl - cf + 1));
                ++oC;
            };
            if (i - f + 1 > 3) {
                int cl_first = f;
                for (int kk = f + 1; kk < i; ++kk) {
                    if (!cut(kk, cl_first, i)) continue;
                    const int cl_last = (kk == i - 1) ? i : kk;
                    emit(cl_first, cl_last);
                    cl_first = kk + 1;
                }
            } else emit(f, i);
        }
        sbase += aS; cbase += aC; ebase += aE;
        if (m) first = t0 + 64 - __clzll(m);
    }
}

__global__ __launch_bounds__(64) void k_classify_root(const MatDesc* __restrict__ md,
                                                      const MatState* __restrict__ ms,
                                                      const BlockRec* __restrict__ blk,
                                                      const int* __restrict__ smat,
                                                      const int* __restrict__ blist,
                                                      const double* __restrict__ gers,
                                                      const double* __restrict__ D,
                                                      const double* __restrict__ E,
                                                      const double* __restrict__ DLr,
                                                      const double* __restrict__ DLLr, double* W,
                                                      const double* __restrict__ Werr,
                                                      const double* __restrict__ Wgap, double* Wsh,
                                                      TreeOut out, const int* __restrict__ nblist_p) {
    const int nblist = *nblist_p;
    const int nblk_ = nblist;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_classify_root_one(bx, md, ms, blk, smat, blist, gers, D, E, DLr, DLLr, W, Werr, Wgap, Wsh, out, nblist);
        __syncthreads();
    }
}

// Deeper levels: one thread per cluster, after its eigenvalues were refined.
__device__ __forceinline__ void k_classify_one(int bx, const ClTask* __restrict__ C, int nC,
                                                 const double* __restrict__ Wgap,
                                                 const double* __restrict__ Wsh, TreeOut out) {
    const int c = bx * blockDim.x + threadIdx.x;
    if (c >= nC || *out.halt) return;
    const ClTask t = C[c];
    classify_cluster(t.cb, t.ce, t.depth, t.mat, t.gblk, t.wseg, t.bbeg, t.bsize, t.bl_spdiam,
                     t.lgap, t.cD, t.cL, t.cDL, t.cDLL, Wgap, Wsh, out);
}

__global__ __launch_bounds__(64) void k_classify(const ClTask* __restrict__ C, const int* __restrict__ nC_p,
                                                 const double* __restrict__ Wgap,
                                                 const double* __restrict__ Wsh, TreeOut out) {
    const int nC = *nC_p;
    const int nblk_ = (nC + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_classify_one(bx, C, nC, Wgap, Wsh, out);
    }
}

// Cuts every block's singletons of this level into wave tasks of at most 64 and
// clears the block counters for the next level.  One thread per block.
// Wave tasks of a level: every block with singletons gets ceil(count/tsize) tasks, listed
// by decreasing block size (size classes of floor(log2 n)) so that the longest tasks
// start first.  k_plan_count sizes the classes, k_plan_s writes the list.
__device__ __forceinline__ int size_class(int n) { return 31 - __clz(max(n, 1)); }

// Clears small counter records with one wave: a one-wave launch fits beside long
// running kernels of other streams, where a fill of several waves may wait for a
// whole compute unit.
__global__ __launch_bounds__(64) void k_zero(int* a, int na, int* b, int nb) {
    for (int i = threadIdx.x; i < na; i += 64) a[i] = 0;
    for (int i = threadIdx.x; i < nb; i += 64) b[i] = 0;
}

// zero_next (optional): the counters of the iteration after this one, which nothing
// reads from here until then; cleared by the first wave.
__device__ __forceinline__ void k_plan_count_one(int bx, const int* __restrict__ blist, int nblist,
                                                   const BlockRec* __restrict__ blk,
                                                   const int* __restrict__ bcnt, int tsize,
                                                   LevelCounters* cnt) {
    const int ib = bx * blockDim.x + threadIdx.x;
    if (ib >= nblist) return;
    const int gb = blist[ib];
    const int c = bcnt[gb];
    if (c == 0) return;
    const int n = blk[gb].bend - blk[gb].bbeg + 1;
    atomicAdd(&cnt->cls[size_class(n)], (c + tsize - 1) / tsize);
}

__global__ __launch_bounds__(64) void k_plan_count(const int* __restrict__ blist, const int* __restrict__ nblist_p,
                                                   const BlockRec* __restrict__ blk,
                                                   const int* __restrict__ bcnt, int tsize,
                                                   LevelCounters* cnt, LevelCounters* zero_next) {
    if (zero_next && blockIdx.x == 0) {
        int* z = reinterpret_cast<int*>(zero_next);
        for (int i = threadIdx.x; i < (int)(sizeof(LevelCounters) / sizeof(int)); i += 64) z[i] = 0;
    }
    const int nblist = *nblist_p;
    const int nblk_ = (nblist + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_plan_count_one(bx, blist, nblist, blk, bcnt, tsize, cnt);
    }
}

__device__ __forceinline__ void k_plan_s_one(int bx, const int* __restrict__ blist, int nblist,
                                               const int* __restrict__ smat,
                                               const MatDesc* __restrict__ md,
                                               const BlockRec* __restrict__ blk, int* bcnt,
                                               int tsize, WaveTask* tasks, int capT,
                                               LevelCounters* cnt) {
    const int ib = bx * blockDim.x + threadIdx.x;
    if (ib >= nblist) return;
    const int gb = blist[ib];
    const int c = bcnt[gb];
    if (c == 0) return;
    bcnt[gb] = 0;
    const int wseg = md[smat[gb]].off + blk[gb].wbeg;
    const int n = blk[gb].bend - blk[gb].bbeg + 1;
    const int q = size_class(n);
    int base = 0;
    for (int k2 = kSizeClasses - 1; k2 > q; --k2) base += cnt->cls[k2];
    const int k = (c + tsize - 1) / tsize;
    base += atomicAdd(&cnt->clsPos[q], k);
    atomicAdd(&cnt->nT, k);
    if (base + k > capT) { atomicOr(&cnt->err, 2); return; }
    for (int t = 0; t < k; ++t) tasks[base + t] = WaveTask{wseg + tsize * t, min(tsize, c - tsize * t)};
}

__global__ __launch_bounds__(64) void k_plan_s(const int* __restrict__ blist, const int* __restrict__ nblist_p,
                                               const int* __restrict__ smat,
                                               const MatDesc* __restrict__ md,
                                               const BlockRec* __restrict__ blk, int* bcnt,
                                               int tsize, WaveTask* tasks, int capT,
                                               LevelCounters* cnt) {
    const int nblist = *nblist_p;
    const int nblk_ = (nblist + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_plan_s_one(bx, blist, nblist, smat, md, blk, bcnt, tsize, tasks, capT, cnt);
    }
}

// Assigns each cluster of a level its slot of 4 * bsize doubles in the solve's pool of
// representations (the slots of all levels of the solve, handed out by one bump
// counter).  A slot that does not fit sets the pool flag and takes the spare slot
// instead (4 * nmax doubles), so that every write stays inside allocated memory: the
// level's classification then emits nothing, the walk ends, and the caller repeats the
// solve with a pool as large as the bump counter grew.
__device__ __forceinline__ void k_arena_assign_one(int bx, ClTask* C, int nC, double* arena,
                                                   unsigned long long* bump,
                                                   unsigned long long cap, double* spare,
                                                   int* flag) {
    const int c = bx * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    ClTask& t = C[c];
    const unsigned long long need = (unsigned long long)(4LL * t.bsize);
    const unsigned long long o = atomicAdd(bump, need);
    double* base;
    if (o + need > cap) {
        atomicOr(flag, 1);
        base = spare;
        t.arena_off = -1;
    } else {
        base = arena + o;
        t.arena_off = (long long)o;
    }
    t.cD = base;
    t.cL = base + t.bsize;
    t.cDL = base + 2 * t.bsize;
    t.cDLL = base + 3 * t.bsize;
}

__global__ __launch_bounds__(64) void k_arena_assign(ClTask* C, const int* __restrict__ nC_p,
                                                     double* arena, unsigned long long* bump,
                                                     unsigned long long cap, double* spare,
                                                     int* flag) {
    const int nC = *nC_p;
    const int nblk_ = (nC + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x)
        k_arena_assign_one(bx, C, nC, arena, bump, cap, spare, flag);
}

// ---------------------------------------------------------------------------------
// k_newrrr: compute_new_rrr (process_cluster.c:105-235) for one cluster at depth >= 1.
// Lanes 0 and 1 refine the two extremal eigenvalues against the parent (the two
// dlarrb calls are independent: each runs with a zero gap); lane 0 then runs dlarrf
// (odrrf.f:140-384); all lanes apply the shift and form DL, DLL.
// ---------------------------------------------------------------------------------
// One dlarrf trial (odrrf.f:206-225 and 231-249): L+ D+ L+^T = L D L^T - sig I, its
// largest |D+| and whether a pivot hit the floor or the growth is NaN.  The loads of
// the next kU steps are issued before the current ones are computed.
template <class P>
__device__ __forceinline__ void rrf_trial(int n, P d, P l, P ld, double sig, double pivmin,
                                          double& mx, int& nan) {
    double s = -sig;
    double dp = d[0] + s;
    nan = 0;
    if (fabs(dp) < pivmin) { dp = -pivmin; nan = 1; }
    mx = fabs(dp);
    double a[kU], b[kU], c[kU], an[kU], bn[kU], cn[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) { a[u] = ld[u]; b[u] = l[u]; c[u] = d[u + 1]; }
    for (int i0 = 0; i0 < n - 1; i0 += kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + kU + u;
            an[u] = ld[i]; bn[u] = l[i]; cn[u] = d[i + 1];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (i0 + u < n - 1) {
                const double lp = a[u] / dp;
                s = s * lp * b[u] - sig;
                dp = c[u] + s;
                if (fabs(dp) < pivmin) { dp = -pivmin; nan = 1; }
                mx = fmax(mx, fabs(dp));
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) { a[u] = an[u]; b[u] = bn[u]; c[u] = cn[u]; }
    }
    nan = nan || isnan(mx);
}

__device__ __forceinline__ void rrf_trial_w(int n, const double* d, const double* l,
                                            const double* ld, double sig, double pivmin,
                                            double& mx, int& nan) {
    if (wave_uniform(d) && wave_uniform(l) && wave_uniform(ld) && wave_uniform(n))
        rrf_trial(rfl(n), as_const(d), as_const(l), as_const(ld), sig, pivmin, mx, nan);
    else
        rrf_trial(n, as_global(d), as_global(l), as_global(ld), sig, pivmin, mx, nan);
}

// The trial factorisation with its factors stored (for the refined growth test); the
// loads run kU steps ahead, as in rrf_trial.
__device__ __forceinline__ void rrf_trial_store(int n, const double* d0, const double* l0,
                                                const double* ld0, double sig, double pivmin,
                                                double* Dst0, double* Lst0) {
    const GPtr d = as_global(d0), l = as_global(l0), ld = as_global(ld0);
    const GPtrW Dst = as_global_w(Dst0), Lst = as_global_w(Lst0);
    double s = -sig;
    double dp = d[0] + s;
    if (fabs(dp) < pivmin) dp = -pivmin;
    Dst[0] = dp;
    double a[kU], b[kU], c[kU], an[kU], bn[kU], cn[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) { a[u] = ld[u]; b[u] = l[u]; c[u] = d[u + 1]; }
    for (int i0 = 0; i0 < n - 1; i0 += kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + kU + u;
            an[u] = ld[i]; bn[u] = l[i]; cn[u] = d[i + 1];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u;
            if (i < n - 1) {
                const double lp = a[u] / dp;
                Lst[i] = lp;
                s = s * lp * b[u] - sig;
                dp = c[u] + s;
                if (fabs(dp) < pivmin) dp = -pivmin;
                Dst[i + 1] = dp;
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) { a[u] = an[u]; b[u] = bn[u]; c[u] = cn[u]; }
    }
}

// The refined growth measure of odrrf.f:265-305 over a D+ array A and an L+ array B
// (the reference pairs the D+ of one trial with the L+ of the other); loads kU ahead.
__device__ __forceinline__ double rrf_rrr(int n, const double* A0, const double* B0, double spdiam) {
    const GPtr A = as_global(A0), B = as_global(B0);
    double tmp = fabs(A[n - 1]), znm2 = 1.0, prod = 1.0, oldp = 1.0;
    double a[kU], b[kU], a1[kU], b1[kU], an[kU], bn[kU];
    // step i reads A[i], B[i] and A[i+1], B[i+1]
#pragma unroll
    for (int u = 0; u < kU; ++u) {
        const int i = n - 2 - u;
        a[u] = A[i]; b[u] = B[i]; a1[u] = A[i + 1]; b1[u] = B[i + 1];
    }
    for (int i0 = n - 2; i0 >= 0; i0 -= kU) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 - kU - u;
            an[u] = A[i]; bn[u] = B[i];
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 - u;
            if (i >= 0) {
                if (prod <= kEps) prod = ((a1[u] * b1[u]) / (a[u] * b[u])) * oldp;
                else prod = prod * fabs(b[u]);
                oldp = prod;
                znm2 = znm2 + prod * prod;
                tmp = fmax(tmp, fabs(a[u] * prod));
            }
        }
        // the next group's A[i+1], B[i+1] are this group's A[i], B[i] shifted by one
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            a1[u] = (u == 0) ? a[kU - 1] : an[u - 1];
            b1[u] = (u == 0) ? b[kU - 1] : bn[u - 1];
            a[u] = an[u];
            b[u] = bn[u];
        }
    }
    return tmp / (spdiam * sqrt(znm2));
}

// The cluster's end values that dlarrf reads (after the ends were refined): W, WERR
// of the first and last eigenvalue, WGAP of the first and of the one before the last,
// and the gap to the right of the cluster.
struct RrfEnds {
    double w0, wl, e0, el, g0, gl, clgapr;
};

// dlarrf (odrrf.f:140-384) for one cluster.  trial(k, side, mx, nan) gives the growth
// and NaN flag of the trial factorisation of try k (0 or 1) at the left (side 0) or
// right (side 1) shift; stored() writes both trial factorisations of the current try
// (left to dplus/lplus, right to work/work + n) for the refined growth test.  The
// trial at the best shift after the last try is accepted without a test, so it is not
// computed (its result is never read).  Returns the shift side; *sigma the shift.
template <class TrialF, class StoreF>
__device__ __forceinline__ int dlarrf_core(int n, const RrfEnds& e, int clsize, double spdiam,
                                           double clgapl, double pivmin, double* sigma,
                                           const double* dplus, const double* lplus,
                                           const double* work, TrialF trial, StoreF stored,
                                           int* sweeps_out) {
    const int KTRYMAX = 1;
    const double fact = (double)(1 << KTRYMAX);
    int shift = 0, forcer = 0, indx = 0;
    const double clwdth = fabs(e.wl - e.w0) + e.el + e.e0;
    const double avgap = clwdth / (double)(clsize - 1);
    const double mingap = fmin(clgapl, e.clgapr);
    double lsigma = fmin(e.w0, e.wl) - e.e0;
    double rsigma = fmax(e.w0, e.wl) + e.el;
    lsigma = lsigma - fabs(lsigma) * 4.0 * kEps;
    rsigma = rsigma + fabs(rsigma) * 4.0 * kEps;
    const double ldmax = 0.25 * mingap + 2.0 * pivmin;
    const double rdmax = 0.25 * mingap + 2.0 * pivmin;
    double ldelta = fmax(avgap, e.g0) / fact;
    double rdelta = fmax(avgap, e.gl) / fact;
    double smlgrowth = 1.0 / kSafmin;
    const double fail2 = (double)(n - 1) * mingap / (spdiam * sqrt(kEps));
    double bestshift = lsigma;
    int ktry = 0;
    const double growthbound = kMaxGrowth1 * spdiam;
    double max1 = 0.0, max2 = 0.0;
    int sweeps = 0;
    for (;;) {
        int sawnan1 = 0, sawnan2 = 0;
        ldelta = fmin(ldmax, ldelta);
        rdelta = fmin(rdmax, rdelta);
        if (forcer) { *sigma = lsigma; shift = 1; break; }
        ++sweeps;
        trial(ktry, 0, lsigma, max1, sawnan1);
        if (max1 <= growthbound && !sawnan1) { *sigma = lsigma; shift = 1; break; }
        ++sweeps;
        trial(ktry, 1, rsigma, max2, sawnan2);
        if (max2 <= growthbound && !sawnan2) { *sigma = rsigma; shift = 2; break; }
        if (!(sawnan1 && sawnan2)) {
            if (!sawnan1) {
                indx = 1;
                if (max1 <= smlgrowth) { smlgrowth = max1; bestshift = lsigma; }
            }
            if (!sawnan2) {
                if (sawnan1 || max2 <= max1) indx = 2;
                if (max2 <= smlgrowth) { smlgrowth = max2; bestshift = rsigma; }
            }
            int dorrr1 = (clwdth < mingap / 128.0) && (fmin(max1, max2) < fail2) && !sawnan1 && !sawnan2;
            if (dorrr1) {
                // the refined test reads both factorisations (odrrf.f:265-305)
                stored(lsigma, rsigma);
                sweeps += 2;
                if (indx == 1) {
                    const double rrr1 = rrf_rrr(n, dplus, work + n, spdiam);
                    if (rrr1 <= kMaxGrowth2) { *sigma = lsigma; shift = 1; break; }
                } else if (indx == 2) {
                    const double rrr2 = rrf_rrr(n, work, lplus, spdiam);
                    if (rrr2 <= kMaxGrowth2) { *sigma = rsigma; shift = 2; break; }
                }
            }
        }
        if (ktry < KTRYMAX) {
            lsigma = fmax(lsigma - ldelta, lsigma - ldmax);
            rsigma = fmin(rsigma + rdelta, rsigma + rdmax);
            ldelta = 2.0 * ldelta;
            rdelta = 2.0 * rdelta;
            ++ktry;
            continue;
        }
        // NOFAIL is .TRUE. in the reference: take the best shift found
        lsigma = bestshift;
        rsigma = bestshift;
        forcer = 1;
    }
    *sweeps_out = sweeps;
    return shift;
}

// The shifts of tries 0 and 1 as dlarrf_core forms them (left and right of each).
__device__ __forceinline__ void rrf_shifts(const RrfEnds& e, int clsize, double clgapl,
                                           double pivmin, double sh[4]) {
    const double fact = 2.0;
    const double clwdth = fabs(e.wl - e.w0) + e.el + e.e0;
    const double avgap = clwdth / (double)(clsize - 1);
    const double mingap = fmin(clgapl, e.clgapr);
    double lsigma = fmin(e.w0, e.wl) - e.e0;
    double rsigma = fmax(e.w0, e.wl) + e.el;
    lsigma = lsigma - fabs(lsigma) * 4.0 * kEps;
    rsigma = rsigma + fabs(rsigma) * 4.0 * kEps;
    const double ldmax = 0.25 * mingap + 2.0 * pivmin;
    const double rdmax = 0.25 * mingap + 2.0 * pivmin;
    const double ldelta = fmin(ldmax, fmax(avgap, e.g0) / fact);
    const double rdelta = fmin(rdmax, fmax(avgap, e.gl) / fact);
    sh[0] = lsigma;
    sh[1] = rsigma;
    sh[2] = fmax(lsigma - ldelta, lsigma - ldmax);
    sh[3] = fmin(rsigma + rdelta, rsigma + rdmax);
}

// The refined ends of cluster t as dlarrf reads them, and their commit
// (process_cluster.c:186-189): WGAP(cb) from the refined left end and the neighbour
// as it was before the right end was refined.
__device__ __forceinline__ RrfEnds rrf_ends(const ClTask& t, const double* Wsh, const double* Werr,
                                            const double* Wgap, const double* Wtmp,
                                            const double* Werrtmp) {
    const int cb = t.cb, ce = t.ce;
    RrfEnds e;
    e.w0 = Wtmp[cb];
    e.e0 = Werrtmp[cb];
    e.wl = Wtmp[ce];
    e.el = Werrtmp[ce];
    e.g0 = fmax(0.0, (Wsh[cb + 1] - Werr[cb + 1]) - (Wtmp[cb] + Werrtmp[cb]));
    e.gl = (ce - 1 == cb) ? e.g0 : Wgap[ce - 1];
    e.clgapr = Wgap[ce];
    return e;
}

// Extremal eigenvalues of every cluster to full accuracy against the parent
// (process_cluster.c:168-195), one lane per cluster end.  The two dlarrb calls of a
// cluster run with a zero gap and do not depend on each other; results go to Wtmp
// and are committed by k_dlarrf.
__device__ __forceinline__ void k_endref_one(int bx, const ClTask* __restrict__ C, int nC,
                                               const int* __restrict__ Windex,
                                               const double* __restrict__ Wsh,
                                               const double* __restrict__ Werr,
                                               const MatState* __restrict__ ms, double* Wtmp,
                                               double* Werrtmp) {
    const int e = bx * blockDim.x + threadIdx.x;
    if (e >= 2 * nC) return;
    const ClTask t = C[e >> 1];
    const int slot = (e & 1) ? t.ce : t.cb;
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    int sw = dlarrb1_w(t.bsize, t.pD, t.pDLL, Windex[slot], S.rqtol, S.rqtol, 0.0, &w, &we,
                       S.pivmin, t.bl_spdiam, t.bsize);
    add_steps(kStepEndref, (unsigned long long)sw * t.bsize);
    Wtmp[slot] = w;
    Werrtmp[slot] = we;
}

__global__ __launch_bounds__(64) void k_endref(const ClTask* __restrict__ C, const int* __restrict__ nC_p,
                                               const int* __restrict__ Windex,
                                               const double* __restrict__ Wsh,
                                               const double* __restrict__ Werr,
                                               const MatState* __restrict__ ms, double* Wtmp,
                                               double* Werrtmp, long long budget) {
    const int nC = *nC_p;
    if (group_lanes_hd(2LL * nC, budget) != 1) return;   // k_endref_grp takes them
    const int nblk_ = (2 * nC + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_endref_one(bx, C, nC, Windex, Wsh, Werr, ms, Wtmp, Werrtmp);
    }
}

// The chosen factorisation L+ D+ L+^T = L D L^T - tau I of each lane's cluster,
// recomputed with the operations of the accepted trial into cD (D+) and cL (L+).
// Each lane writes chunks of 16 values to LDS; the wave then stores 16 contiguous
// rows of four clusters per instruction, so no cache line is left partly written.
// The loads run kU steps ahead of the arithmetic.
template <class P>
__device__ __forceinline__ void rrf_store(int n, int nmax, P d, P l, P ld, double tau,
                                          double pivmin, bool valid, int lane,
                                          double (*bD)[65], double (*bL)[65], double** dD,
                                          double** dL, const int* nn) {
    double s = -tau;
    double dp = d[0] + s;
    if (fabs(dp) < pivmin) dp = -pivmin;
    const int sub = lane & 15, cq = lane >> 4;
    double a[kU], b[kU], c[kU], an[kU], bn[kU], cn[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) { a[u] = ld[u]; b[u] = l[u]; c[u] = d[u + 1]; }
    for (int i0 = 0; i0 < nmax; i0 += 16) {
#pragma unroll
        for (int h = 0; h < 16; h += kU) {
            // the wave runs to its longest block: a lane loads only while its own
            // sweep goes on (then at most kU + 1 entries past its block)
            if (i0 + h + kU < n) {
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const int i = i0 + h + kU + u;
                    an[u] = ld[i]; bn[u] = l[i]; cn[u] = d[i + 1];
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int i = i0 + h + u;
                if (i < n) {
                    bD[h + u][lane] = dp;
                    if (i < n - 1) {
                        const double lp = a[u] / dp;
                        bL[h + u][lane] = lp;
                        s = s * lp * b[u] - tau;
                        dp = c[u] + s;
                        if (fabs(dp) < pivmin) dp = -pivmin;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) { a[u] = an[u]; b[u] = bn[u]; c[u] = cn[u]; }
        }
        __syncthreads();
        for (int cg = 0; cg < 16; ++cg) {
            const int col = cg * 4 + cq;
            const int row = i0 + sub;
            const int nc = nn[col];
            if (row < nc) {
                as_global_w(dD[col])[row] = bD[sub][col];
                if (row < nc - 1) as_global_w(dL[col])[row] = bL[sub][col];
            }
        }
        __syncthreads();
    }
    (void)valid;
}

// The shift of every cluster's new representation by dlarrf (odrrf.f:140-384), with
// G = 1 or 4 lanes per cluster.  With 4, the lanes evaluate the trials at the left and
// right shifts of both tries at once (the shifts do not depend on the trial results),
// and every lane of the group then runs the reference's sequence of tests on them, so
// the shift is the same as with one lane; the refined growth test's two stored trials
// run on two lanes side by side.  The ends refined by k_endref are committed here.
// With one lane per cluster the kernel also stores the chosen factorisation; with four,
// k_rrf_store does.
template <int G>
__device__ __forceinline__ void k_dlarrf_one(int bx, ClTask* C, int nC, double* Wsh, double* Werr,
                                               double* Wgap, const double* __restrict__ Wtmp,
                                               const double* __restrict__ Werrtmp,
                                               const MatState* __restrict__ ms) {
    const int lane = threadIdx.x;
    const int sub = lane % G, gbase = lane - sub;
    const int c = bx * (64 / G) + lane / G;
    const bool valid = c < nC;
    ClTask& t = C[min(c, nC - 1)];
    const int cb = t.cb, ce = t.ce, clsize = ce - cb + 1, n = t.bsize;
    const RrfEnds e = rrf_ends(t, Wsh, Werr, Wgap, Wtmp, Werrtmp);
    const double pivmin = ms[t.mat].pivmin;
    double tau = 0.0;
    int sw = 0, sh;
    if constexpr (G == 1) {
        auto trial = [&](int, int, double sig, double& mx, int& nan) {
            rrf_trial_w(n, t.pD, t.pL, t.pDL, sig, pivmin, mx, nan);
        };
        auto stored = [&](double ls, double rs) {
            if (!valid) return;   // lanes past the last cluster only replay it
            rrf_trial_store(n, t.pD, t.pL, t.pDL, ls, pivmin, t.cD, t.cL);
            rrf_trial_store(n, t.pD, t.pL, t.pDL, rs, pivmin, t.cDL, t.cDL + n);
        };
        sh = dlarrf_core(n, e, clsize, t.bl_spdiam, t.lgap, pivmin, &tau, t.cD, t.cL, t.cDL, trial,
                         stored, &sw);
    } else {
        double shv[4];
        rrf_shifts(e, clsize, t.lgap, pivmin, shv);
        double mx;
        int nn;
        rrf_trial_w(n, t.pD, t.pL, t.pDL, shv[sub & 3], pivmin, mx, nn);
        double pm[4];
        int pn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            pm[k] = __shfl(mx, gbase + k);
            pn[k] = __shfl(nn, gbase + k);
        }
        auto trial = [&](int k, int side, double, double& m2, int& nan2) {
            m2 = pm[2 * k + side];
            nan2 = pn[2 * k + side];
        };
        auto stored = [&](double ls, double rs) {
            if (valid && sub < 2)   // lanes past the last cluster only replay it
                rrf_trial_store(n, t.pD, t.pL, t.pDL, sub == 0 ? ls : rs, pivmin,
                                sub == 0 ? t.cD : t.cDL, sub == 0 ? t.cL : t.cDL + n);
            __threadfence_block();   // the group reads both stored trials next
        };
        sh = dlarrf_core(n, e, clsize, t.bl_spdiam, t.lgap, pivmin, &tau, t.cD, t.cL, t.cDL, trial,
                         stored, &sw);
    }
    if (valid && sub == 0) {
        add_steps(kStepRrf, (unsigned long long)sw * n);
        add_steps(kStepRrfN, (unsigned long long)n);
        Wgap[cb] = e.g0;
        Wsh[cb] = e.w0;
        Werr[cb] = e.e0;
        Wsh[ce] = e.wl;
        Werr[ce] = e.el;
        t.tau = tau;
        t.shift = sh;
    }
    if constexpr (G == 1) {
        // one lane per cluster: the chosen factorisation is stored right away, while
        // the parent representation is still in the caches
        __shared__ double bD[16][65], bL[16][65];
        __shared__ double* dD[64];
        __shared__ double* dL[64];
        __shared__ int nn[64];
        const double *pd = t.pD, *pl = t.pL, *pld = t.pDL;
        dD[lane] = t.cD;
        dL[lane] = t.cL;
        nn[lane] = valid ? n : 0;
        int nmax = valid ? n : 0;
        for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
        const double tv = valid ? tau : 0.0;
        if (wave_uniform(pd) && wave_uniform(pl) && wave_uniform(pld) && wave_uniform(n))
            rrf_store(rfl(n), nmax, as_const(pd), as_const(pl), as_const(pld), tv, pivmin, valid,
                      lane, bD, bL, dD, dL, nn);
        else
            rrf_store(n, nmax, as_global(pd), as_global(pl), as_global(pld), tv, pivmin, valid, lane,
                      bD, bL, dD, dL, nn);
    }
}

template <int G>
__global__ __launch_bounds__(64) void k_dlarrf(ClTask* C, const int* __restrict__ nC_p, int rrf_max, double* Wsh, double* Werr,
                                               double* Wgap, const double* __restrict__ Wtmp,
                                               const double* __restrict__ Werrtmp,
                                               const MatState* __restrict__ ms) {
    const int nC = *nC_p;
    if ((G == 1) != (nC > rrf_max)) return;   // one lane per cluster only past rrf_max
    const int nblk_ = (nC + 64 / G - 1) / (64 / G);
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_dlarrf_one<G>(bx, C, nC, Wsh, Werr, Wgap, Wtmp, Werrtmp, ms);
        __syncthreads();
    }
}

// The chosen factorisation L+ D+ L+^T = L D L^T - tau I of every cluster, recomputed
// with the operations of the accepted trial, into cD (D+) and cL (L+); one lane per
// cluster, stored through LDS (rrf_store).
__device__ __forceinline__ void k_rrf_store_one(int bx, const ClTask* __restrict__ C, int nC,
                                                  const MatState* __restrict__ ms) {
    __shared__ double bD[16][65], bL[16][65];
    __shared__ double* dD[64];
    __shared__ double* dL[64];
    __shared__ int nn[64];
    const int lane = threadIdx.x;
    const int c = bx * blockDim.x + lane;
    const bool valid = c < nC;
    // lanes past the last cluster sweep the last cluster's parent (nothing is stored
    // for them), so every lane reads valid memory
    const ClTask& t = C[min(c, nC - 1)];
    const double tau = valid ? t.tau : 0.0;
    const double pivmin = ms[t.mat].pivmin;
    int n = t.bsize;
    const double *pd = t.pD, *pl = t.pL, *pld = t.pDL;
    dD[lane] = t.cD;
    dL[lane] = t.cL;
    nn[lane] = valid ? n : 0;
    int nmax = valid ? n : 0;
    for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
    if (wave_uniform(pd) && wave_uniform(pl) && wave_uniform(pld) && wave_uniform(n))
        rrf_store(rfl(n), nmax, as_const(pd), as_const(pl), as_const(pld), tau, pivmin, valid, lane,
                  bD, bL, dD, dL, nn);
    else
        rrf_store(n, nmax, as_global(pd), as_global(pl), as_global(pld), tau, pivmin, valid, lane, bD,
                  bL, dD, dL, nn);
}

__global__ __launch_bounds__(64) void k_rrf_store(const ClTask* __restrict__ C, const int* __restrict__ nC_p, int rrf_max,
                                                  const MatState* __restrict__ ms) {
    const int nC = *nC_p;
    if (nC > rrf_max) return;   // k_dlarrf<1> stored them
    const int nblk_ = (nC + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_rrf_store_one(bx, C, nC, ms);
        __syncthreads();
    }
}

// Accumulates the shift, forms DL and DLL from the chosen factorisation and moves
// the cluster's eigenvalues to the new representation (process_cluster.c:206-232).
__device__ __forceinline__ void k_newrrr_fin_one(int bx, const ClTask* __restrict__ C, int nC,
                                                   double* Wsh, double* Werr) {
    const int c = bx;
    if (c >= nC) return;
    const ClTask t = C[c];
    const int n = t.bsize;
    const int lane = threadIdx.x;
    for (int i = lane; i < n; i += 64) {
        const double dv = t.cD[i];
        const double lv = (i < n - 1) ? t.cL[i] : 0.0;
        if (i < n - 1) {
            double tmp = dv * lv;
            t.cDL[i] = tmp;
            t.cDLL[i] = tmp * lv;
        } else {
            t.cL[n - 1] = t.pL[n - 1] + t.tau;
            t.cDL[i] = 0.0;
            t.cDLL[i] = 0.0;
        }
    }
    const double tau = t.tau;
    for (int s = t.cb + lane; s <= t.ce; s += 64) {
        double fudge = 3.0 * DBL_EPSILON * fabs(Wsh[s]);
        Wsh[s] -= tau;
        fudge += 4.0 * DBL_EPSILON * fabs(Wsh[s]);
        Werr[s] += fudge;
    }
}

__global__ __launch_bounds__(64) void k_newrrr_fin(const ClTask* __restrict__ C, const int* __restrict__ nC_p,
                                                   double* Wsh, double* Werr) {
    const int nC = *nC_p;
    const int nblk_ = nC;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_newrrr_fin_one(bx, C, nC, Wsh, Werr);
    }
}

// Refinement of a cluster's eigenvalues against its new representation, one lane
// per eigenvalue (process_cluster.c:414-418; the cluster is one dlarrb call).
__device__ __forceinline__ void k_refine_cl_one(int bx, const ClTask* __restrict__ C,
                                                   const int* __restrict__ eig_slot,
                                                   const int* __restrict__ eig_cl, int nE,
                                                   const int* __restrict__ Windex,
                                                   const double* __restrict__ Wsh,
                                                   const double* __restrict__ Werr,
                                                   const double* __restrict__ Wgap,
                                                   const MatState* __restrict__ ms, double* Wtmp,
                                                   double* Werrtmp) {
    const int e = bx * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int slot = eig_slot[e];
    const ClTask t = C[eig_cl[e]];
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    double rgap = Wgap[slot];
    double lgap = (slot == t.cb) ? rgap : Wgap[slot - 1];
    int sw = dlarrb1_w(t.bsize, t.cD, t.cDLL, Windex[slot], S.rtol1, S.rtol2, fmin(lgap, rgap),
                       &w, &we, S.pivmin, t.bl_spdiam, t.bsize);
    add_steps(kStepRefineCl, (unsigned long long)sw * t.bsize);
    Wtmp[slot] = w;
    Werrtmp[slot] = we;
}

__global__ __launch_bounds__(64) void k_refine_cl(const ClTask* __restrict__ C,
                                                   const int* __restrict__ eig_slot,
                                                   const int* __restrict__ eig_cl, const int* __restrict__ nE_p,
                                                   const int* __restrict__ Windex,
                                                   const double* __restrict__ Wsh,
                                                   const double* __restrict__ Werr,
                                                   const double* __restrict__ Wgap,
                                                   const MatState* __restrict__ ms, double* Wtmp,
                                                   double* Werrtmp, long long budget) {
    const int nE = *nE_p;
    if (group_lanes_hd(nE, budget) != 1) return;   // k_refine_cl_grp takes them
    const int nblk_ = (nE + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_refine_cl_one(bx, C, eig_slot, eig_cl, nE, Windex, Wsh, Werr, Wgap, ms, Wtmp, Werrtmp);
    }
}

__device__ __forceinline__ void k_gaps_cl_one(int bx, const ClTask* __restrict__ C,
                                                 const int* __restrict__ eig_slot,
                                                 const int* __restrict__ eig_cl, int nE,
                                                 const double* __restrict__ Wtmp,
                                                 const double* __restrict__ Werrtmp, double* Wsh,
                                                 double* Werr, double* Wgap, double* W) {
    const int e = bx * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int slot = eig_slot[e];
    const ClTask t = C[eig_cl[e]];
    const double w = Wtmp[slot], we = Werrtmp[slot];
    Wsh[slot] = w;
    Werr[slot] = we;
    if (slot < t.ce) Wgap[slot] = fmax(0.0, Wtmp[slot + 1] - Werrtmp[slot + 1] - w - we);
    W[slot] = w + t.cL[t.bsize - 1];
}

__global__ __launch_bounds__(64) void k_gaps_cl(const ClTask* __restrict__ C,
                                                 const int* __restrict__ eig_slot,
                                                 const int* __restrict__ eig_cl, const int* __restrict__ nE_p,
                                                 const double* __restrict__ Wtmp,
                                                 const double* __restrict__ Werrtmp, double* Wsh,
                                                 double* Werr, double* Wgap, double* W) {
    const int nE = *nE_p;
    const int nblk_ = (nE + 63) / 64;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_gaps_cl_one(bx, C, eig_slot, eig_cl, nE, Wtmp, Werrtmp, Wsh, Werr, Wgap, W);
    }
}

// ---------------------------------------------------------------------------------
// Eigenvectors: dlar1v and its vector solve for a wave of singletons of one block,
// then the singleton kernel (process_singleton.c:65-324; dlar1v odr1v.f:156-369).
// ---------------------------------------------------------------------------------
constexpr int SW = 64;          // lanes per wave
constexpr int kTailN = 2048;    // largest block of the tail kernel (2 (n + 1) doubles of LDS)
constexpr int KC = kChunk;       // steps per chunk: LDS buffer rows and checkpoint spacing
constexpr int kUS = 4;          // prefetch group of the full stationary sweep

__device__ __forceinline__ int wave_max_u(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_min_u(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// Per-lane checkpoints of one wave in device memory, lane-interleaved so that every
// access is one 512-byte row: S(c) is the stationary carry s entering step c*KC,
// P(c) the progressive carry p entering the top step of chunk c (steps c*KC+1 ..
// c*KC+KC).  Only these carries live in device memory; the L+, U-, S and z values
// of a chunk are recomputed from them into the LDS buffer with the same operations,
// so every value equals the one a full-length work array would hold.
struct Ck {
    GPtrW base;
    int nck;
    int lane;
    __device__ __forceinline__ GDouble& S(int c) const { return base[(size_t)c * SW + lane]; }
    __device__ __forceinline__ GDouble& P(int c) const {
        return base[(size_t)(nck + c) * SW + lane];
    }
};

// One column per lane, padded so that the transposed reads of the column writes are
// free of bank conflicts.
typedef double ChunkBuf[KC][SW + 1];

// Destination of the eigenvector columns of a wave (LDS tables filled by the solve).
struct ColOut {
    double** zcol;   // column pointer of each lane
    int* wr0;        // per column: first row of the chunk in the LDS buffer
    int* wlo;        // per column: rows to write, [wlo, whi]
    int* whi;
    double* Zc;      // this lane's column
    int bbeg, nfull, cntw;
};

struct R1v {
    int negcnt, r, isupp0, isupp1, nan1, nan2;
    double ztz, mingma, nrminv, resid, rqcorr;
};

struct Rrr {
    GPtr D;
    GPtr L;
    GPtr LD;
    GPtr LLD;
};

// One step t (1-based) of the stationary transform, odr1v.f:187-191; with safe set
// the form of odr1v.f:206-213 (pivot floor, S = LLD where L+ underflows).
template <class P>
__device__ __forceinline__ void stat_step(double d, double ld, double l, P LLD, int t,
                                          double lambda, bool safe, double pivmin, double& s,
                                          double& lp, double& si, bool& neg) {
    double dplus = d + s;
    if (safe && fabs(dplus) < pivmin) dplus = -pivmin;
    lp = ld / dplus;
    neg = dplus < 0.0;
    si = s * lp * l;
    if (safe && lp == 0.0) si = LLD[t - 1];
    s = si - lambda;
}

// One step t of the progressive transform, odr1v.f:236-241; with safe set the form
// of odr1v.f:247-255.
__device__ __forceinline__ void prog_step(double d, double lld, double l, double lambda, bool safe,
                                          double pivmin, double& p, double& um, bool& neg) {
    double dminus = lld + p;
    if (safe && fabs(dminus) < pivmin) dminus = -pivmin;
    const double tmp = d / dminus;
    neg = dminus < 0.0;
    um = l * tmp;
    double pn = p * tmp - lambda;
    if (safe && tmp == 0.0) pn = d - lambda;
    p = pn;
}

// Stationary transform over steps t = 1 .. stop-1 (odr1v.f:184-229), keeping the
// carry every KC steps.  Lanes with stop <= 1 do nothing.  Outputs the final carry,
// the negative pivots among steps t <= r1-1 and the last S value (0 if no step).
__device__ void pass_stat(const Rrr& R, double lambda, double pivmin, int r1, int stop, bool safe,
                          const Ck& ck, double& s_out, int& neg_out, double& slast_out) {
    double s = 0.0 - lambda, slast = 0.0;
    int neg = 0;
    const int stopU = wave_max_u(stop);
    if (stopU > 1) {
        // two register buffers in turn: the loads of one group are in flight while the
        // other group is computed, and no register copy waits for them
        double A[kUS][3], B[kUS][3];
        auto load = [&](double(&X)[kUS][3], int t0) {
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                const int j = t0 + u - 1;
                X[u][0] = R.D[j]; X[u][1] = R.LD[j]; X[u][2] = R.L[j];
            }
        };
        auto run = [&](const double(&X)[kUS][3], int t0) {
            if (t0 > 0 && (t0 % KC) == 0 && t0 < stop) ck.S(t0 / KC) = s;
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                const int t = t0 + u;
                if (t >= 1 && t < stop) {
                    double lp, si;
                    bool ng;
                    stat_step(X[u][0], X[u][1], X[u][2], R.LLD, t, lambda, safe, pivmin, s, lp, si, ng);
                    neg += (t <= r1 - 1) && ng;
                    slast = si;
                }
            }
        };
        load(A, 0);
        for (int t0 = 0; t0 < stopU; t0 += 2 * kUS) {
            load(B, t0 + kUS);
            run(A, t0);
            if (t0 + kUS >= stopU) break;
            load(A, t0 + 2 * kUS);
            run(B, t0 + kUS);
        }
    }
    s_out = s;
    neg_out = neg;
    slast_out = slast;
}

// Visits steps t = lo .. hi (wave-uniform bounds, ascending) and calls f(t, v) with
// v[a] = P[a][t - 1].  Two register buffers take turns: the loads of one group of kUS
// steps are in flight while the other group is computed.
template <int NA, typename F>
__device__ __forceinline__ void sweep_up(int lo, int hi, const GPtr (&P)[NA], F f) {
    if (lo > hi) return;
    double A[kUS][NA], B[kUS][NA];
    auto load = [&](double(&X)[kUS][NA], int t0) {
#pragma unroll
        for (int u = 0; u < kUS; ++u)
#pragma unroll
            for (int a = 0; a < NA; ++a) X[u][a] = P[a][t0 + u - 1];
    };
    auto run = [&](const double(&X)[kUS][NA], int t0) {
#pragma unroll
        for (int u = 0; u < kUS; ++u)
            if (t0 + u <= hi) f(t0 + u, X[u]);
    };
    load(A, lo);
    for (int t0 = lo; t0 <= hi; t0 += 2 * kUS) {
        load(B, t0 + kUS);
        run(A, t0);
        if (t0 + kUS > hi) break;
        load(A, t0 + 2 * kUS);
        run(B, t0 + kUS);
    }
}

// Descending twin of sweep_up: t = hi down to lo.
template <int NA, typename F>
__device__ __forceinline__ void sweep_down(int hi, int lo, const GPtr (&P)[NA], F f) {
    if (lo > hi) return;
    double A[kUS][NA], B[kUS][NA];
    auto load = [&](double(&X)[kUS][NA], int t0) {
#pragma unroll
        for (int u = 0; u < kUS; ++u)
#pragma unroll
            for (int a = 0; a < NA; ++a) X[u][a] = P[a][t0 - u - 1];
    };
    auto run = [&](const double(&X)[kUS][NA], int t0) {
#pragma unroll
        for (int u = 0; u < kUS; ++u)
            if (t0 - u >= lo) f(t0 - u, X[u]);
    };
    load(A, hi);
    for (int t0 = hi; t0 >= lo; t0 -= 2 * kUS) {
        load(B, t0 - kUS);
        run(A, t0);
        if (t0 - kUS < lo) break;
        load(A, t0 - 2 * kUS);
        run(B, t0 - kUS);
    }
}

// Progressive transform from step n-1 down to r1 with the gamma values and the
// twist index (odr1v.f:231-283).  Free-twist lanes (r1 = 1, r2 = n) scan every gamma;
// the scan runs downwards here, which selects the same index as the upward scan of
// the reference (the last index of minimal |gamma| wins there; NaNs never win except
// at R1).  The S values come from recomputing each chunk into the LDS buffer.
__device__ void pass_prog(int n, const Rrr& R, double lambda, double pivmin, int r1, bool freetw,
                          bool act, bool safe1, bool safe2, double slast, const Ck& ck,
                          ChunkBuf& buf, int lane, double& p_out, int& neg2_out,
                          double& g0raw_out, double& mingma_out, int& r_out) {
    const double pn = R.D[n - 1] - lambda;   // P(BN-1)
    double p = pn;
    int neg2 = 0;
    bool have = false;
    double best = 0.0, g0raw = 0.0;
    int rbest = 0;
    if (act && freetw) {   // gamma at index n-1 (twist n), the first of the downward scan
        double g = slast + pn;
        if (g == 0.0) g = kEps * slast;
        if (!isnan(g)) { have = true; best = g; rbest = n; }
    }
    if (act && !freetw && r1 == n) g0raw = slast + pn;
    const int lo = act ? r1 : n;   // this lane's steps: n-1 down to lo
    const int loU = wave_min_u(lo);
    const bool needS = act && freetw;
    const bool anyfree = __any(needS);
    if (n - 1 >= loU) {
        const GPtr PS[3] = {R.D, R.LD, R.L};
        const GPtr PP[3] = {R.D, R.LLD, R.L};
        const int cTop = (n - 2) / KC, cBot = (loU - 1) / KC;
        double sck = (needS && cTop >= 1) ? ck.S(cTop) : 0.0;
        for (int c = cTop; c >= cBot; --c) {
            const int base = c * KC;
            const int tlo = base + 1, thi = min(base + KC, n - 1);
            if (anyfree) {
                double s = (c == 0) ? 0.0 - lambda : sck;
                sck = (needS && c - 1 >= 1) ? ck.S(c - 1) : 0.0;   // next chunk's carry
                if (c == 0) buf[0][lane] = 0.0;                    // S(0)
                sweep_up<3>(max(base, 1), thi - 1, PS, [&](int j, const double(&v)[3]) {
                    if (needS) {
                        double lp, si;
                        bool ng;
                        stat_step(v[0], v[1], v[2], R.LLD, j, lambda, safe1, pivmin, s, lp, si, ng);
                        buf[j - base][lane] = si;
                    }
                });
            }
            if (act && lo <= thi) ck.P(c) = p;
            sweep_down<3>(thi, max(tlo, loU), PP, [&](int t, const double(&v)[3]) {
                if (act && t >= lo) {
                    double um;
                    bool ng;
                    prog_step(v[0], v[1], v[2], lambda, safe2, pivmin, p, um, ng);
                    neg2 += ng;
                    if (freetw) {
                        const double sg = buf[t - 1 - base][lane];   // S(t-1)
                        double gm = sg + p;
                        if (t == 1) {
                            g0raw = gm;
                        } else {
                            if (gm == 0.0) gm = kEps * sg;
                            if (!isnan(gm) && (!have || fabs(gm) < fabs(best))) {
                                have = true;
                                best = gm;
                                rbest = t;
                            }
                        }
                    } else if (t == r1) {
                        g0raw = slast + p;
                    }
                }
            });
        }
    }
    double mingma;
    int r;
    if (freetw) {
        double g0 = g0raw;
        if (g0 == 0.0) g0 = kEps * 0.0;   // EPS * S(0)
        if (isnan(g0) || !have || fabs(g0) < fabs(best)) { mingma = g0; r = 1; }
        else { mingma = best; r = rbest; }
    } else {
        mingma = g0raw;
        if (fabs(mingma) == 0.0) mingma = kEps * slast;
        r = r1;
    }
    p_out = p;
    neg2_out = neg2;
    g0raw_out = g0raw;
    mingma_out = mingma;
    r_out = r;
}

// dlar1v up to the twist (odr1v.f:156-283) for the lanes with act set; rin is the
// twist index or 0 for a free twist.
__device__ void r1v_ck(int n, double lambda, const Rrr& R, double pivmin, int wantnc, int rin,
                       bool act, const Ck& ck, ChunkBuf& buf, int lane, R1v* o) {
    const int r1 = rin ? rin : 1, r2 = rin ? rin : n;
    const bool freetw = rin == 0;
    double s, slast;
    int neg1;
    pass_stat(R, lambda, pivmin, r1, act ? r2 : 1, false, ck, s, neg1, slast);
    const bool nan1 = act && isnan(s);
    if (__any(nan1)) {   // slow stationary form (odr1v.f:202-228)
        double s2, sl2;
        int ng2;
        pass_stat(R, lambda, pivmin, r1, nan1 ? r2 : 1, true, ck, s2, ng2, sl2);
        if (nan1) { neg1 = ng2; slast = sl2; }
    }
    double p, g0raw, mingma;
    int neg2, r;
    pass_prog(n, R, lambda, pivmin, r1, freetw, act, nan1, false, slast, ck, buf, lane, p, neg2,
              g0raw, mingma, r);
    const bool nan2 = act && isnan(p);
    if (__any(nan2)) {   // slow progressive form (odr1v.f:244-257)
        double p2, g2, m2;
        int ng2, rr2;
        pass_prog(n, R, lambda, pivmin, r1, freetw, nan2, nan1, true, slast, ck, buf, lane, p2,
                  ng2, g2, m2, rr2);
        if (nan2) { neg2 = ng2; g0raw = g2; mingma = m2; r = rr2; }
    }
    if (act) {
        if (g0raw < 0.0) ++neg1;
        o->negcnt = wantnc ? neg1 + neg2 : -1;
        o->r = r;
        o->mingma = mingma;
        o->nan1 = nan1;
        o->nan2 = nan2;
    }
}

// Column write of one chunk with per-lane rows: column c takes rows wr0[c] + sub for
// sub = 0 .. KC-1 within [wlo[c], whi[c]] (1-based block rows) from buf[sub][c].
__device__ __forceinline__ void write_chunk(const ChunkBuf& buf, const ColOut& co, int lane,
                                            int row0, int rlo, int rhi) {
    co.wr0[lane] = row0;
    co.wlo[lane] = rlo;
    co.whi[lane] = rhi;
    __syncthreads();
    // KC rows of SW / KC columns per store instruction
    const int sub = lane % KC, cq = lane / KC;
#pragma unroll 4
    for (int cg = 0; cg < KC; ++cg) {
        const int c = cg * (SW / KC) + cq;
        if (c < co.cntw) {
            const int row = co.wr0[c] + sub;
            if (row >= co.wlo[c] && row <= co.whi[c])
                as_global_w(co.zcol[c])[co.bbeg + row - 1] = buf[sub][c];
        }
    }
    __syncthreads();
}

// The vector solve of odr1v.f:285-355 for the lanes with act set.  Every lane walks its
// own chunks (KC rows each): upwards from the twist with L+ recomputed from the
// stationary checkpoints, downwards with U- recomputed from the progressive ones, and
// stops at its support boundary, so a wave costs its longest support rather than the
// union of its lanes' ranges.  The first loads and the checkpoint of a lane's next
// chunk are issued while the current chunk is computed.  Every active lane forms ZTZ,
// the residual and the support; lanes with stg set also write z * scale into their
// columns (co) for the rows their chunks cover, as the reference's dlar1v writes Z
// on every call (rows outside the final support are cleared when the vector is
// normalised).
__device__ __forceinline__ void pass_solve(int n, const Rrr& R, double lambda, double pivmin,
                                           double gaptol, const Ck& ck, ChunkBuf& buf,
                                           ChunkBuf& bld, int lane, R1v* o, bool act, bool stg,
                                           double scale, const ColOut& co) {
    const int r = act ? o->r : 1;
    const bool safe1 = act && o->nan1, safe2 = act && o->nan2;
    const bool nan = safe1 || safe2;
    int is0 = 1, is1 = n;
    double ztz = 1.0;
    int upper_lo = n + 1, lower_hi = 0;
    const bool stage = __any(act && stg);   // some lane writes its column
    // Waves in which no lane saw a NaN in its transforms (almost all) run the solve
    // without the slow-form branches, each row's updates as selects instead of
    // divergent branches; the others run the reference's slow forms.
    auto solve = [&](auto slow_t) {
        constexpr bool kSlow = decltype(slow_t)::value;
        if constexpr (kSlow) {
            // upwards: Z(j) for j = r-1 .. 1; chunk c holds steps c*KC .. c*KC+KC-1
            {
                double zi1 = 1.0, zi2 = 0.0;   // Z(j+1), Z(j+2)
                // LD(j+1), the LD of the row solved before row j; the first row (j = r-1) has
                // Z(r) = 1 and never reads it
                double ld_next = 0.0;
                bool live = act && r - 1 >= 1;
                int c = live ? (r - 1) / KC : 0;
                double pre[kUS][3];
                double pck = 0.0;
        #pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int j = c * KC + u;
                    pre[u][0] = R.D[j - 1]; pre[u][1] = R.LD[j - 1]; pre[u][2] = R.L[j - 1];
                }
                if (live && c >= 1) pck = ck.S(c);
                while (__any(live)) {
                    const bool mine = live;
                    const int base = c * KC;
                    const int jlo = max(base, 1), jhi = min(base + KC - 1, r - 1);
                    double s = (c == 0) ? 0.0 - lambda : pck;
                    double cur[kUS][3], nxt[kUS][3];
        #pragma unroll
                    for (int u = 0; u < kUS; ++u)
        #pragma unroll
                        for (int a = 0; a < 3; ++a) cur[u][a] = pre[u][a];
                    // next chunk of this lane (c - 1): first loads and carry
                    const int cn = max(c - 1, 0);
        #pragma unroll
                    for (int u = 0; u < kUS; ++u) {
                        const int j = cn * KC + u;
                        pre[u][0] = R.D[j - 1]; pre[u][1] = R.LD[j - 1]; pre[u][2] = R.L[j - 1];
                    }
                    pck = (mine && c - 1 >= 1) ? ck.S(c - 1) : 0.0;
                    // recompute L+ (and keep LD) for the chunk
        #pragma unroll
                    for (int g = 0; g < KC / kUS; ++g) {
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u) {
                                const int j = base + (g + 1) * kUS + u;
                                nxt[u][0] = R.D[j - 1]; nxt[u][1] = R.LD[j - 1]; nxt[u][2] = R.L[j - 1];
                            }
                        }
        #pragma unroll
                        for (int u = 0; u < kUS; ++u) {
                            const int k = g * kUS + u, j = base + k;
                            if (mine && j >= jlo && j <= jhi) {
                                double lp, si;
                                bool ng;
                                stat_step(cur[u][0], cur[u][1], cur[u][2], R.LLD, j, lambda, safe1, pivmin,
                                          s, lp, si, ng);
                                buf[k][lane] = lp;
                                bld[k][lane] = cur[u][1];
                            }
                        }
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u)
        #pragma unroll
                                for (int a = 0; a < 3; ++a) cur[u][a] = nxt[u][a];
                        }
                    }
                    // consume downwards
        #pragma unroll
                    for (int k = KC - 1; k >= 0; --k) {
                        const int j = base + k;
                        if (mine && j >= jlo && j <= jhi) {
                            const double ldj = bld[k][lane];
                            double zv = 0.0;
                            if (live) {
                                double zi;
                                if (nan && zi1 == 0.0) zi = -(ld_next / ldj) * zi2;
                                else zi = -(buf[k][lane] * zi1);
                                ld_next = ldj;
                                if ((fabs(zi) + fabs(zi1)) * fabs(ldj) < gaptol) {
                                    is0 = j + 1;
                                    live = false;
                                } else {
                                    ztz = ztz + zi * zi;
                                    zv = zi;
                                    zi2 = zi1;
                                    zi1 = zi;
                                }
                            }
                            if (stg) buf[k][lane] = zv * scale;
                        }
                    }
                    if (mine) upper_lo = jlo;
                    if (stage) write_chunk(buf, co, lane, base, mine && stg ? jlo : 1, mine && stg ? jhi : 0);
                    if (c == 0) live = false;
                    c = cn;
                }
            }

            // downwards: Z(t+1) for t = r .. n-1; chunk c holds steps c*KC+1 .. c*KC+KC
            {
                double zi = 1.0, zim1 = 0.0;   // Z(t), Z(t-1)
                // LD(t-1), the LD of the row solved before row t; the first row (t = r) has
                // Z(r) = 1 and never reads it
                double ld_prev = 0.0;
                bool live = act && r <= n - 1;
                const int cEnd = (n - 2) / KC;
                int c = live ? (r - 1) / KC : 0;
                double pre[kUS][4];
                double pcp = 0.0;
        #pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int t = c * KC + KC - u;
                    pre[u][0] = R.D[t - 1]; pre[u][1] = R.LLD[t - 1]; pre[u][2] = R.L[t - 1];
                    pre[u][3] = R.LD[t - 1];
                }
                if (live) pcp = ck.P(c);
                while (__any(live)) {
                    const bool mine = live;
                    const int tlo = c * KC + 1, thi = min(c * KC + KC, n - 1);
                    const int tl2 = max(tlo, r);
                    double p = pcp;
                    double cur[kUS][4], nxt[kUS][4];
        #pragma unroll
                    for (int u = 0; u < kUS; ++u)
        #pragma unroll
                        for (int a = 0; a < 4; ++a) cur[u][a] = pre[u][a];
                    const int cn = min(c + 1, cEnd);
        #pragma unroll
                    for (int u = 0; u < kUS; ++u) {
                        const int t = cn * KC + KC - u;
                        pre[u][0] = R.D[t - 1]; pre[u][1] = R.LLD[t - 1]; pre[u][2] = R.L[t - 1];
                        pre[u][3] = R.LD[t - 1];
                    }
                    pcp = (mine && c + 1 <= cEnd) ? ck.P(c + 1) : 0.0;
                    // recompute U- (and keep LD) for the chunk, from its top down
        #pragma unroll
                    for (int g = 0; g < KC / kUS; ++g) {
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u) {
                                const int t = tlo + KC - 1 - (g + 1) * kUS - u;
                                nxt[u][0] = R.D[t - 1]; nxt[u][1] = R.LLD[t - 1]; nxt[u][2] = R.L[t - 1];
                                nxt[u][3] = R.LD[t - 1];
                            }
                        }
        #pragma unroll
                        for (int u = 0; u < kUS; ++u) {
                            const int k = KC - 1 - (g * kUS + u), t = tlo + k;
                            if (mine && t >= tl2 && t <= thi) {
                                double um;
                                bool ng;
                                prog_step(cur[u][0], cur[u][1], cur[u][2], lambda, safe2, pivmin, p, um, ng);
                                buf[k][lane] = um;
                                bld[k][lane] = cur[u][3];
                            }
                        }
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u)
        #pragma unroll
                                for (int a = 0; a < 4; ++a) cur[u][a] = nxt[u][a];
                        }
                    }
                    // consume upwards in t
        #pragma unroll
                    for (int k = 0; k < KC; ++k) {
                        const int t = tlo + k;
                        if (mine && t >= tl2 && t <= thi) {
                            const double ldt = bld[k][lane];
                            double zv = 0.0;
                            if (live) {
                                double zn;
                                if (nan && zi == 0.0) zn = -(ld_prev / ldt) * zim1;
                                else zn = -(buf[k][lane] * zi);
                                ld_prev = ldt;
                                if ((fabs(zi) + fabs(zn)) * fabs(ldt) < gaptol) {
                                    is1 = t;
                                    live = false;
                                } else {
                                    ztz = ztz + zn * zn;
                                    zv = zn;
                                    zim1 = zi;
                                    zi = zn;
                                }
                            }
                            if (stg) buf[k][lane] = zv * scale;
                        }
                    }
                    if (mine) lower_hi = thi + 1;
                    if (stage)
                        write_chunk(buf, co, lane, tlo + 1, mine && stg ? tl2 + 1 : 1, mine && stg ? thi + 1 : 0);
                    if (c == cEnd) live = false;
                    c = cn;
                }
            }

        } else {
            // upwards: Z(j) for j = r-1 .. 1; chunk c holds steps c*KC .. c*KC+KC-1
            {
                double zi1 = 1.0;   // Z(j+1)
                bool live = act && r - 1 >= 1;
                int c = live ? (r - 1) / KC : 0;
                double pre[kUS][3];
                double pck = 0.0;
        #pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int j = c * KC + u;
                    pre[u][0] = R.D[j - 1]; pre[u][1] = R.LD[j - 1]; pre[u][2] = R.L[j - 1];
                }
                if (live && c >= 1) pck = ck.S(c);
                while (__any(live)) {
                    const bool mine = live;
                    const int base = c * KC;
                    const int jlo = max(base, 1), jhi = min(base + KC - 1, r - 1);
                    double s = (c == 0) ? 0.0 - lambda : pck;
                    double cur[kUS][3], nxt[kUS][3];
        #pragma unroll
                    for (int u = 0; u < kUS; ++u)
        #pragma unroll
                        for (int a = 0; a < 3; ++a) cur[u][a] = pre[u][a];
                    // next chunk of this lane (c - 1): first loads and carry
                    const int cn = max(c - 1, 0);
        #pragma unroll
                    for (int u = 0; u < kUS; ++u) {
                        const int j = cn * KC + u;
                        pre[u][0] = R.D[j - 1]; pre[u][1] = R.LD[j - 1]; pre[u][2] = R.L[j - 1];
                    }
                    pck = (mine && c - 1 >= 1) ? ck.S(c - 1) : 0.0;
                    // recompute L+ (and keep LD) for the chunk
        #pragma unroll
                    for (int g = 0; g < KC / kUS; ++g) {
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u) {
                                const int j = base + (g + 1) * kUS + u;
                                nxt[u][0] = R.D[j - 1]; nxt[u][1] = R.LD[j - 1]; nxt[u][2] = R.L[j - 1];
                            }
                        }
        #pragma unroll
                        for (int u = 0; u < kUS; ++u) {
                            const int k = g * kUS + u, j = base + k;
                            const bool in = mine && j >= jlo && j <= jhi;
                            double sn = s, lp, si;
                            bool ng;
                            stat_step(cur[u][0], cur[u][1], cur[u][2], R.LLD, j, lambda, false, pivmin, sn,
                                      lp, si, ng);
                            s = in ? sn : s;
                            buf[k][lane] = lp;
                            bld[k][lane] = cur[u][1];
                        }
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u)
        #pragma unroll
                                for (int a = 0; a < 3; ++a) cur[u][a] = nxt[u][a];
                        }
                    }
                    // consume downwards
        #pragma unroll
                    for (int k = KC - 1; k >= 0; --k) {
                        const int j = base + k;
                        const bool in = mine && j >= jlo && j <= jhi;
                        const double ldj = bld[k][lane];
                        const double zi = -(buf[k][lane] * zi1);
                        const bool go = in && live;
                        const bool cut = go && (fabs(zi) + fabs(zi1)) * fabs(ldj) < gaptol;
                        const bool upd = go && !cut;
                        is0 = cut ? j + 1 : is0;
                        live = cut ? false : live;
                        ztz = upd ? ztz + zi * zi : ztz;
                        zi1 = upd ? zi : zi1;
                        if (stage) buf[k][lane] = (upd ? zi : 0.0) * scale;
                    }
                    if (mine) upper_lo = jlo;
                    if (stage) write_chunk(buf, co, lane, base, mine && stg ? jlo : 1, mine && stg ? jhi : 0);
                    if (c == 0) live = false;
                    c = cn;
                }
            }

            // downwards: Z(t+1) for t = r .. n-1; chunk c holds steps c*KC+1 .. c*KC+KC
            {
                double zi = 1.0;   // Z(t)
                bool live = act && r <= n - 1;
                const int cEnd = (n - 2) / KC;
                int c = live ? (r - 1) / KC : 0;
                double pre[kUS][4];
                double pcp = 0.0;
        #pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int t = c * KC + KC - u;
                    pre[u][0] = R.D[t - 1]; pre[u][1] = R.LLD[t - 1]; pre[u][2] = R.L[t - 1];
                    pre[u][3] = R.LD[t - 1];
                }
                if (live) pcp = ck.P(c);
                while (__any(live)) {
                    const bool mine = live;
                    const int tlo = c * KC + 1, thi = min(c * KC + KC, n - 1);
                    const int tl2 = max(tlo, r);
                    double p = pcp;
                    double cur[kUS][4], nxt[kUS][4];
        #pragma unroll
                    for (int u = 0; u < kUS; ++u)
        #pragma unroll
                        for (int a = 0; a < 4; ++a) cur[u][a] = pre[u][a];
                    const int cn = min(c + 1, cEnd);
        #pragma unroll
                    for (int u = 0; u < kUS; ++u) {
                        const int t = cn * KC + KC - u;
                        pre[u][0] = R.D[t - 1]; pre[u][1] = R.LLD[t - 1]; pre[u][2] = R.L[t - 1];
                        pre[u][3] = R.LD[t - 1];
                    }
                    pcp = (mine && c + 1 <= cEnd) ? ck.P(c + 1) : 0.0;
                    // recompute U- (and keep LD) for the chunk, from its top down
        #pragma unroll
                    for (int g = 0; g < KC / kUS; ++g) {
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u) {
                                const int t = tlo + KC - 1 - (g + 1) * kUS - u;
                                nxt[u][0] = R.D[t - 1]; nxt[u][1] = R.LLD[t - 1]; nxt[u][2] = R.L[t - 1];
                                nxt[u][3] = R.LD[t - 1];
                            }
                        }
        #pragma unroll
                        for (int u = 0; u < kUS; ++u) {
                            const int k = KC - 1 - (g * kUS + u), t = tlo + k;
                            const bool in = mine && t >= tl2 && t <= thi;
                            double pn = p, um;
                            bool ng;
                            prog_step(cur[u][0], cur[u][1], cur[u][2], lambda, false, pivmin, pn, um, ng);
                            p = in ? pn : p;
                            buf[k][lane] = um;
                            bld[k][lane] = cur[u][3];
                        }
                        if (g + 1 < KC / kUS) {
        #pragma unroll
                            for (int u = 0; u < kUS; ++u)
        #pragma unroll
                                for (int a = 0; a < 4; ++a) cur[u][a] = nxt[u][a];
                        }
                    }
                    // consume upwards in t
        #pragma unroll
                    for (int k = 0; k < KC; ++k) {
                        const int t = tlo + k;
                        const bool in = mine && t >= tl2 && t <= thi;
                        const double ldt = bld[k][lane];
                        const double zn = -(buf[k][lane] * zi);
                        const bool go = in && live;
                        const bool cut = go && (fabs(zi) + fabs(zn)) * fabs(ldt) < gaptol;
                        const bool upd = go && !cut;
                        is1 = cut ? t : is1;
                        live = cut ? false : live;
                        ztz = upd ? ztz + zn * zn : ztz;
                        zi = upd ? zn : zi;
                        if (stage) buf[k][lane] = (upd ? zn : 0.0) * scale;
                    }
                    if (mine) lower_hi = thi + 1;
                    if (stage)
                        write_chunk(buf, co, lane, tlo + 1, mine && stg ? tl2 + 1 : 1, mine && stg ? thi + 1 : 0);
                    if (c == cEnd) live = false;
                    c = cn;
                }
            }

        }
    };
    if (__any(nan)) solve(std::true_type{});
    else solve(std::false_type{});
    if (act && stg) as_global_w(co.Zc)[co.bbeg + r - 1] = scale;   // Z(r) = 1
    (void)upper_lo;
    (void)lower_hi;
    if (act) {
        const double tmp = 1.0 / ztz;
        o->ztz = ztz;
        o->nrminv = sqrt(tmp);
        o->resid = fabs(o->mingma) * o->nrminv;
        o->rqcorr = o->mingma * tmp;
        o->isupp0 = is0;
        o->isupp1 = is1;
    }
}

// ---------------------------------------------------------------------------------
// Eigenvectors of singletons (PMR_process_s_task, process_singleton.c:65-324): one
// lane per eigenvector, run as bulk iterations.  Each launch of k_s_iter performs
// one Rayleigh-quotient iteration (dlar1v to the twist, the vector solve and the
// update of odrrv.f / process_singleton.c) for every singleton still iterating,
// packed 64 to a wave by block; k_s_collect compacts the ones that go on, and the
// bisection fallbacks run in their own packed launch (k_s_bisect), so a lane that
// needs many iterations or a bisection no longer holds a wave of finished ones.
// Every solve writes its unscaled vector into the column, as the reference's dlar1v
// writes Z on every call; k_s_normalize scales the final support by 1/||z|| and
// clears the rest of the column.  The per-singleton state between launches is an
// SState record.
// ---------------------------------------------------------------------------------
enum : int { kSActive = 0, kSStep2 = 1, kSDone = 2 };
__device__ __forceinline__ int sflag_state(int f) { return (f >> 4) & 3; }

// The Rayleigh-quotient step after a dlar1v of singleton idx (process_singleton.c:
// 185-263): convergence test, RQ correction or bisection request, best residual, and
// on completion the support, the eigenvalue and its error.  Stores the new state.
__device__ __forceinline__ void s_update(SState& q, const R1v& o, bool s2, double lambda,
                                         int usedBS, int usedRQ, int needBS, int wantnc,
                                         int state, int i_local, double tol, double gap,
                                         const MatDesc& dsc, int col, int bbeg, const STask& t,
                                         int n, double* W, double* Wsh, double* Werr, SState* ST,
                                         int idx) {
    const double RQtol = 2.0 * DBL_EPSILON;
    bool done = false;
    q.lamv = lambda;
    if (s2) {
        done = true;
    } else {
        int r = o.r;
        double left = q.left, right = q.right, bstres = q.bstres, bstw = q.bstw;
        const int kk = q.kk;
        bool conv = false;
        if (kk == 1 || o.resid < bstres) { bstres = o.resid; bstw = lambda; }
        if (!usedBS && o.resid > tol * gap && fabs(o.rqcorr) > RQtol * fabs(lambda)) {
            int sgndef = (i_local <= o.negcnt) ? -1 : 1;
            if (o.rqcorr * sgndef >= 0.0 && lambda + o.rqcorr <= right &&
                lambda + o.rqcorr >= left) {
                usedRQ = 1;
                if (sgndef == 1) left = lambda;
                else right = lambda;
                q.wsh = 0.5 * (left + right);
                lambda += o.rqcorr;
            } else needBS = 1;
            if (right - left < RQtol * fabs(lambda)) usedBS = 1;
            else if (kk == kMaxIterRQI - 1) needBS = 1;
        } else conv = true;
        q.kk = kk + 1;
        q.r = r;
        q.left = left; q.right = right; q.bstres = bstres; q.bstw = bstw;
        q.lam = lambda;
        if (conv || q.kk > kMaxIterRQI) {
            // the best-residual step after a bisection, else the vector is final
            if (usedRQ && usedBS && bstres <= o.resid) state = kSStep2;
            else done = true;
        }
    }
    if (done) {
        state = kSDone;
        as_global_w(dsc.Zsupp)[2 * col] = o.isupp0 + bbeg;
        as_global_w(dsc.Zsupp)[2 * col + 1] = o.isupp1 + bbeg;
        Wsh[t.slot] = lambda;
        Werr[t.slot] = q.werr;
        W[t.slot] = lambda + t.L[n - 1];
        q.lam = lambda;
    }
    q.resid = o.resid;
    q.nrminv = o.nrminv;
    q.is0 = o.isupp0;
    q.is1 = o.isupp1;
    q.flags = usedBS | (usedRQ << 1) | (needBS << 2) | (wantnc << 3) | (state << 4);
    ST[idx] = q;
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_s_iter(
    const WaveTask* __restrict__ tasks, LevelCounters* cnt, int first, const int* __restrict__ SI,
    const STask* __restrict__ T, const MatDesc* __restrict__ md, MatState* __restrict__ ms,
    const int* __restrict__ Windex, const int* __restrict__ Zidx, double* W, double* Wsh,
    double* Werr, const double* __restrict__ Wgap, SState* ST, double* scratch, int cap,
    int tailS, int tailN) {
    __shared__ ChunkBuf buf, bld;
    __shared__ double* zcol[SW];
    __shared__ int wr0[SW], wlo[SW], whi[SW];
    __shared__ int s_task;
    const int lane = threadIdx.x;
    const int nT = __builtin_amdgcn_readfirstlane(cnt->nT);
    // late iterations with few singletons leave those of blocks up to tailN rows to the
    // tail kernels (all blocks when the large-block tail kernel runs)
    const bool leave_small = !first && __builtin_amdgcn_readfirstlane(cnt->nS) <= tailS;
    Ck ck;
    ck.nck = cap / KC + 2;
    ck.lane = lane;
    ck.base = as_global_w(scratch + (size_t)blockIdx.x * 2 * (size_t)ck.nck * SW);
    for (;;) {
        // tasks are handed out in list order (longest blocks first) as waves free up;
        // the index goes through LDS between barriers so that every lane of the wave
        // takes the same task (a value read from one lane of a lane-0-only branch is
        // not safe against the compiler splitting the wave)
        __syncthreads();
        if (lane == 0) s_task = atomicAdd(&cnt->next, 1);
        __syncthreads();
        const int ti = s_task;
        if (ti >= nT) break;
        const unsigned long long tick0 = wall_clock64();
        const WaveTask wt = tasks[ti];
        const int cntw = __builtin_amdgcn_readfirstlane(wt.cnt);
        const int start = __builtin_amdgcn_readfirstlane(wt.start);
        const bool valid = lane < cntw;
        // lanes past the end replay the last entry so that every pointer is valid
        const int pos = start + min(lane, cntw - 1);
        const int idx = first ? pos : SI[pos];
        const STask t = T[idx];
        const int n = __builtin_amdgcn_readfirstlane(t.bsize);
        if (leave_small && n <= tailN) continue;
        const int bbeg = __builtin_amdgcn_readfirstlane(t.bbeg);
        const MatDesc& dsc = md[t.mat];
        const int col = Zidx[t.slot];
        if (__any(col >= dsc.mcols)) {
            // no column past the caller's Z: k_zindex fails such a matrix before its
            // tasks exist, and should one appear all the same, the wave writes nothing
            // and fails every matrix it holds
            ms[t.mat].status = kTooFew;
            continue;
        }
        ColOut co;
        co.zcol = zcol;
        co.wr0 = wr0;
        co.wlo = wlo;
        co.whi = whi;
        co.Zc = dsc.Z + (size_t)col * dsc.ldz;
        co.bbeg = bbeg;
        co.nfull = __builtin_amdgcn_readfirstlane(dsc.n);
        co.cntw = cntw;
        zcol[lane] = co.Zc;
        if (n == 1) {
            // a 1x1 block: e_bbeg, finished at once
            if (valid) {
                co.Zc[bbeg] = 1.0;
                dsc.Zsupp[2 * col] = 1 + bbeg;
                dsc.Zsupp[2 * col + 1] = 1 + bbeg;
                SState q;
                q.lam = 0.0; q.lamv = 0.0; q.werr = 0.0; q.left = 0.0; q.right = 0.0;
                q.wsh = 0.0; q.bstres = 0.0; q.bstw = 0.0; q.nrminv = 1.0; q.resid = 0.0;
                q.kk = 0; q.r = 1; q.flags = kSDone << 4; q.is0 = 1; q.is1 = 1; q.pad = 0;
                ST[idx] = q;
            }
            continue;
        }
        Rrr R;
        R.D = as_global(t.D); R.L = as_global(t.L); R.LD = as_global(t.DL); R.LLD = as_global(t.DLL);
        const MatState& S = ms[t.mat];
        const int i = t.slot - dsc.off;   // matrix-local eigenvalue index
        const int bl_begin = bbeg, bl_end = bbeg + n - 1;
        const double pivmin = S.pivmin;
        const double tol = 4.0 * log((double)n) * DBL_EPSILON;
        // gaps from the eigenvalue as the level handed it over (Wsh/Werr of the slot
        // change only when the vector is done)
        const double w0 = Wsh[t.slot], e0 = Werr[t.slot];
        const double lg = (i == bl_begin) ? DBL_EPSILON * fmax(fabs(w0 - e0), fabs(w0 + e0)) : t.lgap;
        const double rg = (i == bl_end) ? DBL_EPSILON * fmax(fabs(w0 - e0), fabs(w0 + e0)) : Wgap[t.slot];
        const double gap = fmin(lg, rg);
        const double gaptol = (i == bl_begin || i == bl_end) ? 0.0 : gap * DBL_EPSILON;
        const int i_local = Windex[t.slot];
        SState q;
        if (first) {
            q.lam = w0; q.lamv = w0; q.werr = e0; q.left = w0 - e0; q.right = w0 + e0;
            q.wsh = w0; q.bstres = 0.0; q.bstw = 0.0; q.nrminv = 1.0; q.resid = 0.0;
            q.kk = 1; q.r = 0; q.flags = (1 << 3) | (kSActive << 4); q.is0 = 1; q.is1 = n; q.pad = 0;
        } else {
            q = ST[idx];
        }
        int usedBS = q.flags & 1, usedRQ = (q.flags >> 1) & 1, needBS = (q.flags >> 2) & 1;
        int wantnc = (q.flags >> 3) & 1, state = sflag_state(q.flags);
        const bool act = valid && (state == kSActive || state == kSStep2);
        const bool s2 = state == kSStep2;
        double lambda = q.lam;
        if (s2) lambda = q.bstw;
        if (act && !s2) wantnc = usedBS ? 0 : 1;
        R1v o;
        o.r = 1; o.nan1 = 0; o.nan2 = 0; o.resid = q.resid; o.rqcorr = 0.0; o.nrminv = q.nrminv;
        o.mingma = 0.0; o.isupp0 = q.is0; o.isupp1 = q.is1; o.negcnt = 0; o.ztz = 1.0;
        const unsigned long long tk0 = wall_clock64();
        r1v_ck(n, lambda, R, pivmin, wantnc, q.r, act, ck, buf, lane, &o);
        const unsigned long long tk1 = wall_clock64();
        // The first solve of a level almost never gives the final vector (nearly every
        // singleton takes a second dlar1v), so it does not write its column; a singleton
        // that finishes here solves once more and writes (the same operations on the
        // same checkpoints give the same z).
        bool sa = act, sw = act && !first;
        for (int pass = 0; pass < 2; ++pass) {
            pass_solve(n, R, lambda, pivmin, gaptol, ck, buf, bld, lane, &o, sa, sw, 1.0, co);
            if (pass == 1) break;
            if (act)
                s_update(q, o, s2, lambda, usedBS, usedRQ, needBS, wantnc, state, i_local, tol, gap,
                         dsc, col, bbeg, t, n, W, Wsh, Werr, ST, idx);
            if (!first) break;
            sa = sw = act && sflag_state(q.flags) == kSDone;
            if (!__any(sa)) break;
        }
        const unsigned long long tk2 = wall_clock64();
        add_steps(kStepR1v, act ? (unsigned long long)n : 0ULL);
        if (lane == 0) {
            add_steps(kStepR1vWave, (unsigned long long)n);
            add_steps(kStepTR1v, tk1 - tk0);
            add_steps(kStepTSolve, tk2 - tk1);
            const unsigned long long tick1 = wall_clock64();
            add_steps(kStepTaskTicks, tick1 - tick0);
            add_steps(kStepTasks, 1);
        }
    }
}

// dlarrb (odrrb.f:135-299) for one eigenvalue by a group of G lanes (G a power of
// two, 4 or more; the group is lanes gbase .. gbase + G - 1 of the wave and every lane
// of it passes the same arguments).  The bracket expansion runs both ends at once: each
// half of the group evaluates G/2 consecutive candidates of the reference's sequence
// for its end (left - back or right + back, back doubled) per sweep and keeps the first
// that brackets; each bisection round evaluates the G - 1 midpoints
// of the next log2(G) halvings of the interval (a binary tree, computed with the
// reference's operations) in one sweep, then walks the tree with the reference's
// convergence test.  The walk takes the same path and the same midpoints as one-point
// bisection, so w and werr are bit-identical to dlarrb1.  count(x) is the negcount
// at x; lanes with valid false replay a valid group and their results are not used.
// Returns the reference's one-point sweeps; *rounds_out gets the group's sweeps.
template <class CountF>
__device__ __forceinline__ long long dlarrb_grp(int G, int sub, int gbase, bool valid,
                                                CountF count, int i, double rtol1,
                                                double rtol2, double gap, double pivmin,
                                                double spdiam, double& w, double& werr,
                                                int* rounds_out) {
    const int maxitr = (int)((log(spdiam + pivmin) - log(pivmin)) / log(2.0)) + 2;
    const double mnwdth = 2.0 * pivmin;
    int rounds = 0;
    // bracket expansion to the left and to the right (odrrb.f:176-196), both in the same
    // sweeps: the lower half of the group walks the left sequence, the upper half the
    // right one, H candidates per sweep each
    const int H = G / 2;
    const int side = sub / H, hs = sub % H, hbase = gbase + side * H;
    double x = side == 0 ? w - werr : w + werr, back = werr;
    long long ssweeps = 0;   // this side's one-point sweeps of the reference
    int done = 0, g0 = 0;
    while (__any(valid && !done)) {
        // candidate hs of this batch: hs more expansion steps from x
        double xs = x, bs = back;
        for (int k = 0; k < hs; ++k) {
            if (side == 0) xs = xs - bs;
            else xs = xs + bs;
            bs = 2.0 * bs;
        }
        const int c = count(xs);
        ++rounds;
        const bool brk = side == 0 ? !(c > i - 1) : !(c < i);
        const bool ok = g0 + hs < kMaxExpand;   // the reference stops after kMaxExpand checks
        const unsigned long long m = __ballot(brk || !ok);
        const unsigned long long hm = (m >> hbase) & ((1ULL << H) - 1ULL);
        if (!done) {   // uniform over the half
            if (hm) {
                const int k = __ffsll((long long)hm) - 1;   // first candidate that brackets (or the cap)
                ssweeps += min(k + 1, kMaxExpand - g0);
                x = __shfl(xs, hbase + k);
                done = 1;
            } else {
                ssweeps += H;
                x = __shfl(xs, hbase + H - 1);
                back = __shfl(bs, hbase + H - 1);
                // advance one more step past the last candidate
                if (side == 0) x = x - back;
                else x = x + back;
                back = 2.0 * back;
                g0 += H;
            }
        }
    }
    double left = __shfl(x, gbase), right = __shfl(x, gbase + H);
    long long sweeps = __shfl(ssweeps, gbase) + __shfl(ssweeps, gbase + H);
    double width = 0.5 * fabs(left - right);
    double tmp = fmax(fabs(left), fabs(right));
    double cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
    int conv = (width <= cvrgd || width <= mnwdth);   // interval already tight: unchanged
    int iter = 0;
    int depth = 0;
    for (int g2 = G; g2 > 1; g2 >>= 1) ++depth;
    bool fin = conv;
    while (__any(valid && !fin)) {
        // midpoint of heap node sub (1 .. G-1) of the next depth halvings
        double xm = 0.0;
        if (sub >= 1) {
            int h = 0;
            while ((sub >> (h + 1)) != 0) ++h;   // depth of node sub
            double a = left, bb = right;
            for (int k = h - 1; k >= 0; --k) {
                const double m2 = 0.5 * (a + bb);
                if ((sub >> k) & 1) a = m2;
                else bb = m2;
            }
            xm = 0.5 * (a + bb);
        }
        const int c = (sub >= 1) ? count(xm) : 0;
        ++rounds;
        // walk the tree with the reference's test at every node
        int j = 1;
        for (int lev = 0; lev < depth; ++lev) {
            const int cj = __shfl(c, gbase + j);
            if (!fin) {
                const double mid = 0.5 * (left + right);
                width = right - mid;
                tmp = fmax(fabs(left), fabs(right));
                cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
                if (width <= cvrgd || width <= mnwdth || iter == maxitr) {
                    fin = true;
                } else {
                    ++sweeps;
                    ++iter;
                    if (cj <= i - 1) { left = mid; j = 2 * j + 1; }
                    else { right = mid; j = 2 * j; }
                }
            }
            if (j >= G) j = G - 1;   // keeps the shuffle source in the group once finished
        }
    }
    if (!conv) {
        w = 0.5 * (left + right);
        werr = right - w;
    }
    *rounds_out = rounds;
    return sweeps;
}

// The cluster refinements of levels with few clusters run with a group of G lanes
// per eigenvalue (G = 4 .. 64, chosen on the host from the level's counts): a lane
// per eigenvalue would leave them a long chain of dependent sweeps on few waves, and
// they gate the next level.  Same results as k_endref and k_refine_cl.
__device__ __forceinline__ void k_endref_grp_one(int bx, const ClTask* __restrict__ C, int nC, int G,
                                                   const int* __restrict__ Windex,
                                                   const double* __restrict__ Wsh,
                                                   const double* __restrict__ Werr,
                                                   const MatState* __restrict__ ms, double* Wtmp,
                                                   double* Werrtmp) {
    const int lane = threadIdx.x;
    const int sub = lane % G, gbase = lane - sub;
    const int ne = 2 * nC;
    const int e = bx * (SW / G) + lane / G;
    const bool valid = e < ne;
    const int ee = min(e, ne - 1);
    const ClTask t = C[ee >> 1];
    const int slot = (ee & 1) ? t.ce : t.cb;
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    const int n = t.bsize;
    const double* d = t.pD;
    const double* lld = t.pDLL;
    const bool uni = wave_uniform(d) && wave_uniform(lld) && wave_uniform(n);
    auto count = [&](double x) -> int {
        return uni ? dlaneg(rfl(n), as_const(d), as_const(lld), x, rfl(n))
                   : dlaneg(n, as_global(d), as_global(lld), x, n);
    };
    int rounds = 0;
    const long long sw = dlarrb_grp(G, sub, gbase, valid, count, Windex[slot], S.rqtol, S.rqtol,
                                    0.0, S.pivmin, t.bl_spdiam, w, we, &rounds);
    if (valid && sub == 0) {
        add_steps(kStepEndref, (unsigned long long)sw * n);
        Wtmp[slot] = w;
        Werrtmp[slot] = we;
    }
}

__global__ __launch_bounds__(64) void k_endref_grp(const ClTask* __restrict__ C, const int* __restrict__ nC_p, long long budget,
                                                   const int* __restrict__ Windex,
                                                   const double* __restrict__ Wsh,
                                                   const double* __restrict__ Werr,
                                                   const MatState* __restrict__ ms, double* Wtmp,
                                                   double* Werrtmp) {
    const int nC = *nC_p;
    const int G = group_lanes_hd(2LL * nC, budget);
    if (G == 1) return;
    const int nblk_ = (2 * nC + SW / G - 1) / (SW / G);
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_endref_grp_one(bx, C, nC, G, Windex, Wsh, Werr, ms, Wtmp, Werrtmp);
    }
}

__device__ __forceinline__ void k_refine_cl_grp_one(int bx, const ClTask* __restrict__ C,
                                                      const int* __restrict__ eig_slot,
                                                      const int* __restrict__ eig_cl, int nE,
                                                      int G, const int* __restrict__ Windex,
                                                      const double* __restrict__ Wsh,
                                                      const double* __restrict__ Werr,
                                                      const double* __restrict__ Wgap,
                                                      const MatState* __restrict__ ms,
                                                      double* Wtmp, double* Werrtmp) {
    const int lane = threadIdx.x;
    const int sub = lane % G, gbase = lane - sub;
    const int e = bx * (SW / G) + lane / G;
    const bool valid = e < nE;
    const int ee = min(e, nE - 1);
    const int slot = eig_slot[ee];
    const ClTask t = C[eig_cl[ee]];
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    const double rgap = Wgap[slot];
    const double lgap = (slot == t.cb) ? rgap : Wgap[slot - 1];
    const int n = t.bsize;
    const double* d = t.cD;
    const double* lld = t.cDLL;
    const bool uni = wave_uniform(d) && wave_uniform(lld) && wave_uniform(n);
    auto count = [&](double x) -> int {
        return uni ? dlaneg(rfl(n), as_const(d), as_const(lld), x, rfl(n))
                   : dlaneg(n, as_global(d), as_global(lld), x, n);
    };
    int rounds = 0;
    const long long sw = dlarrb_grp(G, sub, gbase, valid, count, Windex[slot], S.rtol1, S.rtol2,
                                    fmin(lgap, rgap), S.pivmin, t.bl_spdiam, w, we, &rounds);
    if (valid && sub == 0) {
        add_steps(kStepRefineCl, (unsigned long long)sw * n);
        Wtmp[slot] = w;
        Werrtmp[slot] = we;
    }
}

__global__ __launch_bounds__(64) void k_refine_cl_grp(const ClTask* __restrict__ C,
                                                      const int* __restrict__ eig_slot,
                                                      const int* __restrict__ eig_cl, const int* __restrict__ nE_p,
                                                      long long budget, const int* __restrict__ Windex,
                                                      const double* __restrict__ Wsh,
                                                      const double* __restrict__ Werr,
                                                      const double* __restrict__ Wgap,
                                                      const MatState* __restrict__ ms,
                                                      double* Wtmp, double* Werrtmp) {
    const int nE = *nE_p;
    const int G = group_lanes_hd(nE, budget);
    if (G == 1) return;
    const int nblk_ = (nE + SW / G - 1) / (SW / G);
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_refine_cl_grp_one(bx, C, eig_slot, eig_cl, nE, G, Windex, Wsh, Werr, Wgap, ms, Wtmp, Werrtmp);
    }
}

// The bisection fallback of the singletons that need it in this iteration
// (process_singleton.c via dlarrb, odrrb.f:135-299) by multisection: a group of G
// lanes (G a power of two) serves one singleton.  The bracket expansion evaluates G
// consecutive candidates of the reference's sequence (left - back, back doubled) per
// sweep and keeps the first that brackets; each bisection round evaluates the G - 1
// midpoints of the next log2(G) halvings of the interval (a binary tree, computed
// with the reference's operations) in one sweep, then walks the tree with the
// reference's convergence test.  The walk takes the same path and the same midpoints
// as one-point bisection, so w and werr are bit-identical to dlarrb.
__device__ __forceinline__ void s_bisect_one(int b, int nb, int G, const int* __restrict__ BI,
                                             const STask* __restrict__ T,
                                             const MatState* __restrict__ ms,
                                             const int* __restrict__ Windex, SState* ST);

// Grid-stride over the bisection list; the group size G is chosen from the list
// length on the device: as many lanes per singleton as keep about 128K lanes busy.
__global__ __launch_bounds__(64) void k_s_bisect(const int* __restrict__ BI,
                                                 const LevelCounters* __restrict__ cnt,
                                                 const STask* __restrict__ T,
                                                 const MatState* __restrict__ ms,
                                                 const int* __restrict__ Windex, SState* ST) {
    const int nb = __builtin_amdgcn_readfirstlane(cnt->nEig);
    if (nb == 0) return;
    int G = 64;
    while (G > 2 && (long long)nb * G > 131072) G >>= 1;
    const int per = SW / G;
    for (int b0 = blockIdx.x * per; b0 < nb; b0 += gridDim.x * per)
        s_bisect_one(b0 + (int)threadIdx.x / G, nb, G, BI, T, ms, Windex, ST);
}

// The bisection fallback of one singleton on a group of G lanes of this wave (lanes
// gbase .. gbase + G - 1, sub = lane - gbase; every lane of the group passes the same
// singleton): dlarrb to full accuracy from the current bracket (process_singleton.c:
// 186-204), then the state the next dlar1v starts from.  Returns the reference's
// one-point sweeps; *rounds gets the group's sweeps.
__device__ __forceinline__ long long s_bisect_state(SState& q, const STask& t,
                                                    const MatState& S, int i, int G, int sub,
                                                    int gbase, bool valid, int* rounds,
                                                    bool one_singleton = false) {
    const int n = t.bsize;
    const double pivmin = S.pivmin;
    const double spdiam = t.bl_spdiam;
    int r = q.r;
    if (r < 1 || r > n) r = n;
    const double* d = t.D;
    const double* lld = t.DLL;
    double w = q.wsh, we = q.werr;
    long long sweeps;
    if (one_singleton) {   // the whole wave serves this singleton: scalar loads only
        auto count = [&](double x) -> int {
            return dlaneg(rfl(n), as_const(d), as_const(lld), x, rfl(r));
        };
        sweeps = dlarrb_grp(G, sub, gbase, valid, count, i, 0.0, 2.0 * DBL_EPSILON, 0.0, pivmin,
                            spdiam, w, we, rounds);
    } else {
        const bool uni = wave_uniform(d) && wave_uniform(lld) && wave_uniform(n) && wave_uniform(r);
        auto count = [&](double x) -> int {
            return uni ? dlaneg(rfl(n), as_const(d), as_const(lld), x, rfl(r))
                       : dlaneg(n, as_global(d), as_global(lld), x, r);
        };
        sweeps = dlarrb_grp(G, sub, gbase, valid, count, i, 0.0, 2.0 * DBL_EPSILON, 0.0, pivmin,
                            spdiam, w, we, rounds);
    }
    q.wsh = w;
    q.werr = we;
    q.lam = w;
    q.r = 0;
    q.flags |= 1;   // usedBS
    return sweeps;
}

__device__ __forceinline__ void s_bisect_one(int b, int nb, int G, const int* __restrict__ BI,
                                             const STask* __restrict__ T,
                                             const MatState* __restrict__ ms,
                                             const int* __restrict__ Windex, SState* ST) {
    const int lane = threadIdx.x;
    const int grp = lane / G, sub = lane % G;
    const bool valid = b < nb;
    const int idx = BI[min(b, nb - 1)];
    const STask t = T[idx];
    SState q = ST[idx];
    int rounds = 0;
    const long long sweeps = s_bisect_state(q, t, ms[t.mat], Windex[t.slot], G, sub, grp * G, valid,
                                            &rounds);
    if (valid && sub == 0) {
        add_steps(kStepR1vBis, (unsigned long long)t.bsize * (unsigned long long)sweeps);
        ST[idx] = q;
    }
    if (lane == 0) add_steps(kStepBisWave, (unsigned long long)t.bsize * rounds);
}

// ---------------------------------------------------------------------------------
// The tail of a level's singleton iterations: once few singletons go on (at most
// tailS, a launch argument), each one of a block up to kTailN rows gets a workgroup of
// two waves, which runs all the iterations left to it in the bulk schedule one after
// the other (with the bisection fallback on its first wave when an iteration asks for
// it), so the level's tail costs the chain of its slowest singleton instead of one
// round of launches per iteration.  The
// stationary transform (wave 0) and the progressive transform (wave 1) run at the
// same time, each on one lane, keeping their S and P values in LDS; the twist search
// is a reduction over all lanes; L+ and U- are recomputed from the stored carries by
// all lanes (each value needs only its own carry); the two halves of the vector solve
// run on one lane of each wave again.  Every operation and its order are those of
// r1v_ck and pass_solve (odr1v.f:156-369), so the results are bit-identical; only
// the lanes that carry them differ.
// ---------------------------------------------------------------------------------

// Prefetch depth of the one-lane sweeps of the tail kernel (scalar loads; a deeper
// buffer spills scalar registers).
constexpr int kUT = 4;

// Stationary transform over steps 1 .. stop-1 (odr1v.f:184-229) on one lane, with S
// after step t in SV[t] (SV[0] = 0).
template <class P, class SP>
__device__ __forceinline__ void tail_stat(P D, P LD, P L, const double* LLD, double lambda,
                                          double pivmin, int r1, int stop, bool safe,
                                          SP SV, double& s_out, int& neg_out,
                                          double& slast_out) {
    double s = 0.0 - lambda;
    int neg = 0;
    SV[0] = 0.0;
    double a[kUT], b[kUT], c[kUT], an[kUT], bn[kUT], cn[kUT];
#pragma unroll
    for (int u = 0; u < kUT; ++u) { a[u] = D[u]; b[u] = LD[u]; c[u] = L[u]; }
    for (int t0 = 1; t0 < stop; t0 += kUT) {
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int j = t0 - 1 + kUT + u;
            an[u] = D[j]; bn[u] = LD[j]; cn[u] = L[j];
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int t = t0 + u;
            if (t < stop) {
                double lp, si;
                bool ng;
                stat_step(a[u], b[u], c[u], LLD, t, lambda, safe, pivmin, s, lp, si, ng);
                neg += (t <= r1 - 1) && ng;
                SV[t] = si;
            }
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) { a[u] = an[u]; b[u] = bn[u]; c[u] = cn[u]; }
    }
    s_out = s;
    neg_out = neg;
    slast_out = stop > 1 ? SV[stop - 1] : 0.0;
}

// Progressive transform from step n-1 down to lo (odr1v.f:231-257) on one lane, with P
// after step t in PV[t] and PV[n] = D(n) - lambda.
template <class P, class SP>
__device__ __forceinline__ void tail_prog(int n, P D, P LLD, P L, double lambda, double pivmin,
                                          int lo, bool safe, SP PV, double& p_out,
                                          int& neg_out) {
    double p = D[n - 1] - lambda;
    PV[n] = p;
    int neg = 0;
    double a[kUT], b[kUT], c[kUT], an[kUT], bn[kUT], cn[kUT];
#pragma unroll
    for (int u = 0; u < kUT; ++u) {
        const int j = n - 2 - u;
        a[u] = D[j]; b[u] = LLD[j]; c[u] = L[j];
    }
    for (int t0 = n - 1; t0 >= lo; t0 -= kUT) {
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int j = t0 - kUT - u - 1;
            an[u] = D[j]; bn[u] = LLD[j]; cn[u] = L[j];
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int t = t0 - u;
            if (t >= lo) {
                double um;
                bool ng;
                prog_step(a[u], b[u], c[u], lambda, safe, pivmin, p, um, ng);
                neg += ng;
                PV[t] = p;
            }
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) { a[u] = an[u]; b[u] = bn[u]; c[u] = cn[u]; }
    }
    p_out = p;
    neg_out = neg;
}

// Upward half of the vector solve (odr1v.f:285-320) on one lane: Z(j) for j = r-1 down
// to the support cut, read L+(j) from SV[j] and written there.  Returns ISUPPZ(1).
template <class P, class SP>
__device__ __forceinline__ int tail_solve_up(int r, P LD, bool nan, double gaptol, SP SV) {
    double zi1 = 1.0, zi2 = 0.0;
    int is0 = 1;
    bool live = r - 1 >= 1;
    double a[kUT], an[kUT];
#pragma unroll
    for (int u = 0; u < kUT; ++u) a[u] = LD[r - 2 - u];
    for (int j0 = r - 1; j0 >= 1 && live; j0 -= kUT) {
#pragma unroll
        for (int u = 0; u < kUT; ++u) an[u] = LD[j0 - kUT - u - 1];
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int j = j0 - u;
            if (live && j >= 1) {
                const double ldj = a[u];
                double zi;
                if (nan && zi1 == 0.0) zi = -(LD[j] / ldj) * zi2;
                else zi = -(SV[j] * zi1);
                if ((fabs(zi) + fabs(zi1)) * fabs(ldj) < gaptol) {
                    is0 = j + 1;
                    live = false;
                } else {
                    SV[j] = zi;
                    zi2 = zi1;
                    zi1 = zi;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) a[u] = an[u];
    }
    return is0;
}

// Downward half (odr1v.f:321-355) on one lane: Z(t+1) for t = r .. n-1 up to the cut,
// read U-(t) from PV[t] and written there.  Returns ISUPPZ(2).
template <class P, class SP>
__device__ __forceinline__ int tail_solve_down(int n, int r, P LD, bool nan, double gaptol,
                                               SP PV) {
    double zi = 1.0, zim1 = 0.0;
    int is1 = n;
    bool live = r <= n - 1;
    double a[kUT], an[kUT];
#pragma unroll
    for (int u = 0; u < kUT; ++u) a[u] = LD[r - 1 + u];
    for (int t0 = r; t0 <= n - 1 && live; t0 += kUT) {
#pragma unroll
        for (int u = 0; u < kUT; ++u) an[u] = LD[t0 + kUT + u - 1];
#pragma unroll
        for (int u = 0; u < kUT; ++u) {
            const int t = t0 + u;
            if (live && t <= n - 1) {
                const double ldt = a[u];
                double zn;
                if (nan && zi == 0.0) zn = -(LD[max(t - 2, 0)] / ldt) * zim1;
                else zn = -(PV[t] * zi);
                if ((fabs(zi) + fabs(zn)) * fabs(ldt) < gaptol) {
                    is1 = t;
                    live = false;
                } else {
                    PV[t] = zn;
                    zim1 = zi;
                    zi = zn;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kUT; ++u) a[u] = an[u];
    }
    return is1;
}

// Twist candidate order: smaller |gamma| first, then the larger index (the index the
// reference's scan keeps among equal minima).
__device__ __forceinline__ bool tw_better(bool h1, double g1, int t1, bool h2, double g2, int t2) {
    if (!h1) return false;
    if (!h2) return true;
    const double a1 = fabs(g1), a2 = fabs(g2);
    return a1 < a2 || (a1 == a2 && t1 > t2);
}

// BIG: blocks above kTailN rows, with S and P in a slice of device memory per
// workgroup (gsv + blockIdx.x * gstride, two arrays of gstride / 2) instead of LDS.
template <bool BIG>
__global__ __launch_bounds__(128) void k_s_tail(const int* __restrict__ DI,
                                                LevelCounters* cnt,
                                                const STask* __restrict__ T,
                                                const MatDesc* __restrict__ md,
                                                MatState* __restrict__ ms,
                                                const int* __restrict__ Windex,
                                                const int* __restrict__ Zidx, double* W,
                                                double* Wsh, double* Werr,
                                                const double* __restrict__ Wgap, SState* ST,
                                                int iters, int tailS, double* gsv, int gstride) {
    __shared__ double SVs[BIG ? 1 : kTailN + 1];   // S values, then L+, then z above the twist
    __shared__ double PVs[BIG ? 1 : kTailN + 1];   // P values, then U-, then z below the twist (shifted)
    const auto SV = [&] {
        if constexpr (BIG) return as_global_w(gsv + (size_t)blockIdx.x * gstride);
        else return &SVs[0];
    }();
    const auto PV = [&] {
        if constexpr (BIG) return as_global_w(gsv + (size_t)blockIdx.x * gstride + gstride / 2);
        else return &PVs[0];
    }();
    __shared__ double sh_slast;
    __shared__ int sh_i[8];             // nan1, neg1, nan2, neg2, is0, is1
    __shared__ double tw_g[2];
    __shared__ int tw_t[2], tw_h[2];
    __shared__ SState sq;               // the singleton's state between its iterations
    __shared__ int s_k;
    if (cnt->nS > tailS) return;
    const int nD = BIG ? cnt->nDb : cnt->nD;
    const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
    for (;;) {
        // singletons are handed out one at a time as workgroups free up (their chains
        // differ by up to ten iterations and a bisection)
        __syncthreads();
        if (tid == 0) s_k = atomicAdd(BIG ? &cnt->nextDb : &cnt->nextD, 1);
        __syncthreads();
        const int k = s_k;
        if (k >= nD) break;
        const int idx = DI[k];
        const STask t = T[idx];
        const int n = t.bsize;
        SState q = ST[idx];
        if (n < 2) continue;   // uniform
        const MatDesc& dsc = md[t.mat];
        const MatState& S = ms[t.mat];
        const int col = Zidx[t.slot];
        if (col >= dsc.mcols) {
            // never past the caller's Z (k_zindex rules it out): write nothing, fail
            if (tid == 0) ms[t.mat].status = kTooFew;
            continue;
        }
        const int bbeg = t.bbeg;
        const int i = t.slot - dsc.off;
        const int bl_begin = bbeg, bl_end = bbeg + n - 1;
        const double pivmin = S.pivmin;
        const double tol = 4.0 * log((double)n) * DBL_EPSILON;
        const double w0 = Wsh[t.slot], e0 = Werr[t.slot];
        const double lg = (i == bl_begin) ? DBL_EPSILON * fmax(fabs(w0 - e0), fabs(w0 + e0)) : t.lgap;
        const double rg = (i == bl_end) ? DBL_EPSILON * fmax(fabs(w0 - e0), fabs(w0 + e0)) : Wgap[t.slot];
        const double gap = fmin(lg, rg);
        const double gaptol = (i == bl_begin || i == bl_end) ? 0.0 : gap * DBL_EPSILON;
        const int i_local = Windex[t.slot];
        // the iterations left to this singleton in the level's bulk schedule, one after
        // the other (the bisection of the first one has run in k_s_bisect)
        for (int itk = 0; itk < iters; ++itk) {
            int usedBS = q.flags & 1, usedRQ = (q.flags >> 1) & 1, needBS = (q.flags >> 2) & 1;
            int wantnc = (q.flags >> 3) & 1, state = sflag_state(q.flags);
            if (!(state == kSActive || state == kSStep2)) break;   // uniform
            if (itk > 0 && state == kSActive && needBS) {
                // the bisection fallback (process_singleton.c:186-204) on wave 0
                if (wv == 0) {
                    int rounds = 0;
                    const long long sw = s_bisect_state(q, t, S, i_local, SW, lane, 0, true, &rounds, true);
                    if (lane == 0) {
                        add_steps(BIG ? kStepTailBBis : kStepTailBis, (unsigned long long)n * (unsigned long long)sw);
                        add_steps(kStepFallbacks, 1);
                        sq = q;
                    }
                }
                __syncthreads();
                q = sq;
                __syncthreads();
                usedBS = q.flags & 1;
            }
            const bool s2 = state == kSStep2;
            double lambda = s2 ? q.bstw : q.lam;
            if (!s2) wantnc = usedBS ? 0 : 1;
            const int rin = q.r;
            const int r1 = rin ? rin : 1, r2 = rin ? rin : n;
            const bool freetw = rin == 0;
            // the two transforms, one lane of each wave
            if (wv == 0 && lane == 0) {
                double sfin, slast;
                int neg1;
                tail_stat(as_const(t.D), as_const(t.DL), as_const(t.L), t.DLL, lambda, pivmin, r1, r2,
                          false, SV, sfin, neg1, slast);
                const int nan1 = isnan(sfin);
                if (nan1)   // slow stationary form (odr1v.f:202-228)
                    tail_stat(as_const(t.D), as_const(t.DL), as_const(t.L), t.DLL, lambda, pivmin, r1,
                              r2, true, SV, sfin, neg1, slast);
                sh_i[0] = nan1;
                sh_i[1] = neg1;
                sh_slast = slast;
            }
            if (wv == 1 && lane == 0) {
                double pfin;
                int neg2;
                tail_prog(n, as_const(t.D), as_const(t.DLL), as_const(t.L), lambda, pivmin, r1, false,
                          PV, pfin, neg2);
                const int nan2 = isnan(pfin);
                if (nan2)   // slow progressive form (odr1v.f:244-257)
                    tail_prog(n, as_const(t.D), as_const(t.DLL), as_const(t.L), lambda, pivmin, r1,
                              true, PV, pfin, neg2);
                sh_i[2] = nan2;
                sh_i[3] = neg2;
            }
            __syncthreads();
            const bool nan1 = sh_i[0], nan2 = sh_i[2];
            const double slast = sh_slast;
            // twist (odr1v.f:259-283): gamma(t) = S(t-1) + P(t) for t = 2 .. n
            double mingma, g0raw;
            int r;
            if (freetw) {
                bool have = false;
                double best = 0.0;
                int rb = 0;
                for (int tt = 2 + tid; tt <= n; tt += 128) {
                    const double sg = SV[tt - 1];
                    double gm = sg + PV[tt];
                    if (gm == 0.0) gm = kEps * sg;
                    if (!isnan(gm) && tw_better(true, gm, tt, have, best, rb)) {
                        have = true;
                        best = gm;
                        rb = tt;
                    }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    const bool h2 = __shfl_xor((int)have, o);
                    const double g2 = __shfl_xor(best, o);
                    const int t2 = __shfl_xor(rb, o);
                    if (tw_better(h2, g2, t2, have, best, rb)) { have = true; best = g2; rb = t2; }
                }
                if (lane == 0) { tw_h[wv] = have; tw_g[wv] = best; tw_t[wv] = rb; }
                __syncthreads();
                have = tw_h[0];
                best = tw_g[0];
                rb = tw_t[0];
                if (tw_better(tw_h[1], tw_g[1], tw_t[1], have, best, rb)) {
                    have = true;
                    best = tw_g[1];
                    rb = tw_t[1];
                }
                g0raw = 0.0 + PV[1];   // gamma(1) = S(0) + P(1)
                double g0 = g0raw;
                if (g0 == 0.0) g0 = kEps * 0.0;
                if (isnan(g0) || !have || fabs(g0) < fabs(best)) { mingma = g0; r = 1; }
                else { mingma = best; r = rb; }
            } else {
                g0raw = slast + PV[r1];
                mingma = g0raw;
                if (fabs(mingma) == 0.0) mingma = kEps * slast;
                r = r1;
            }
            const int negcnt = wantnc ? sh_i[1] + (g0raw < 0.0 ? 1 : 0) + sh_i[3] : -1;
            // L+(j) for j < r and U-(t) for t >= r, each from its own carry, in place: L+
            // in blocks of 128 rows from the top of its range down and U- from the bottom
            // of its range up, so that no block overwrites a carry a later block reads
            {
                const int nl = r - 1, nu = n - r;   // rows 1 .. r-1 and r .. n-1
                const int nb = (max(nl, nu) + 127) / 128;
                for (int bk = 0; bk < nb; ++bk) {
                    const int j = r - 1 - 128 * bk - tid;   // L+ row
                    const int u = r + 128 * bk + tid;       // U- row
                    double lv = 0.0, uv = 0.0;
                    if (j >= 1 && j <= r - 1) {
                        double dplus = t.D[j - 1] + (SV[j - 1] - lambda);
                        if (nan1 && fabs(dplus) < pivmin) dplus = -pivmin;
                        lv = t.DL[j - 1] / dplus;
                    }
                    if (u >= r && u <= n - 1) {
                        double dminus = t.DLL[u - 1] + PV[u + 1];
                        if (nan2 && fabs(dminus) < pivmin) dminus = -pivmin;
                        uv = t.L[u - 1] * (t.D[u - 1] / dminus);
                    }
                    __syncthreads();
                    if (j >= 1 && j <= r - 1) SV[j] = lv;
                    if (u >= r && u <= n - 1) PV[u] = uv;
                    __syncthreads();
                }
            }
            // the two halves of the solve, one lane of each wave
            const bool nan = nan1 || nan2;
            if (wv == 0 && lane == 0) sh_i[4] = tail_solve_up(r, as_const(t.DL), nan, gaptol, SV);
            if (wv == 1 && lane == 0) sh_i[5] = tail_solve_down(n, r, as_const(t.DL), nan, gaptol, PV);
            __syncthreads();
            const int is0 = sh_i[4], is1 = sh_i[5];
            // the column: z for the rows of the support, Z(r) = 1
            double* Zc = dsc.Z + (size_t)col * dsc.ldz + bbeg - 1;
            for (int row = is0 + tid; row <= is1; row += 128)
                Zc[row] = row < r ? SV[row] : (row == r ? 1.0 : PV[row - 1]);
            if (tid == 0) {
                // ZTZ in the reference's order: the upper half from the twist up, then the
                // lower half from the twist down
                double ztz = 1.0;
                for (int j = r - 1; j >= is0; --j) ztz = ztz + SV[j] * SV[j];
                for (int tt = r; tt <= is1 - 1; ++tt) ztz = ztz + PV[tt] * PV[tt];
                R1v o;
                o.negcnt = negcnt;
                o.r = r;
                o.nan1 = nan1;
                o.nan2 = nan2;
                o.mingma = mingma;
                const double tmp = 1.0 / ztz;
                o.ztz = ztz;
                o.nrminv = sqrt(tmp);
                o.resid = fabs(mingma) * o.nrminv;
                o.rqcorr = mingma * tmp;
                o.isupp0 = is0;
                o.isupp1 = is1;
                s_update(q, o, s2, lambda, usedBS, usedRQ, needBS, wantnc, state, i_local, tol, gap, dsc,
                         col, bbeg, t, n, W, Wsh, Werr, ST, idx);
                add_steps(BIG ? kStepTailBR1v : kStepTailR1v, (unsigned long long)n);
                sq = q;
            }
            __syncthreads();
            q = sq;
            __syncthreads();
        }
    }
}

// Collects the singletons that go on iterating (per block, segmented like the
// singleton list; also in the dense list of the tail kernel: DI for blocks up to kTailN
// rows, DIb for larger ones) and the ones that need the bisection fallback first.  One wave per
// task of the previous iteration.  Counters: nS = singletons going on, nEig = of
// them needing bisection.
__device__ __forceinline__ void k_s_collect_one(int bx, const WaveTask* __restrict__ tasks,
                                                  const LevelCounters* __restrict__ pcnt,
                                                  int first, const int* __restrict__ SIp,
                                                  const STask* __restrict__ T,
                                                  const SState* __restrict__ ST, int* SIn,
                                                  int* bcnt, int* BI, int* DI, int* DIb,
                                                  LevelCounters* cnt) {
    if ((int)bx >= pcnt->nT) return;
    const WaveTask wt = tasks[bx];
    const int lane = threadIdx.x;
    const bool valid = lane < wt.cnt;
    const int idx = valid ? (first ? wt.start + lane : SIp[wt.start + lane]) : 0;
    const int f = valid ? ST[idx].flags : (kSDone << 4);
    const int st = sflag_state(f);
    const bool on = valid && st != kSDone;
    // one atomic per wave and list (a wave's task is one block, so the block counter is
    // wave-uniform too); each lane takes its position from the ballot below it
    const unsigned long long below = (1ULL << lane) - 1ULL;
    const unsigned long long m_on = __ballot(on);
    if (m_on == 0) return;
    const int lead = __ffsll((long long)m_on) - 1;
    const STask& t = T[on ? idx : 0];
    const bool small = on && t.bsize <= kTailN;
    const bool big = on && t.bsize > kTailN;
    const bool bis = on && st == kSActive && ((f >> 2) & 1);
    const unsigned long long m_s = __ballot(small), m_b = __ballot(big), m_i = __ballot(bis);
    int b_on = 0, b_s = 0, b_b = 0, b_i = 0;
    const int gblk = __shfl(on ? t.gblk : 0, lead);
    const int wseg = __shfl(on ? t.wseg : 0, lead);
    if (lane == lead) {
        b_on = atomicAdd(&bcnt[gblk], __popcll(m_on));
        atomicAdd(&cnt->nS, __popcll(m_on));
        if (m_s) b_s = atomicAdd(&cnt->nD, __popcll(m_s));
        if (m_b) b_b = atomicAdd(&cnt->nDb, __popcll(m_b));
        if (m_i) {
            b_i = atomicAdd(&cnt->nEig, __popcll(m_i));
            add_steps(kStepFallbacks, (unsigned long long)__popcll(m_i));
        }
    }
    b_on = __shfl(b_on, lead);
    b_s = __shfl(b_s, lead);
    b_b = __shfl(b_b, lead);
    b_i = __shfl(b_i, lead);
    if (on) SIn[wseg + b_on + __popcll(m_on & below)] = idx;
    if (small) DI[b_s + __popcll(m_s & below)] = idx;
    if (big) DIb[b_b + __popcll(m_b & below)] = idx;
    if (bis) BI[b_i + __popcll(m_i & below)] = idx;
}

__global__ __launch_bounds__(64) void k_s_collect(const WaveTask* __restrict__ tasks,
                                                  const LevelCounters* __restrict__ pcnt,
                                                  int first, const int* __restrict__ SIp,
                                                  const STask* __restrict__ T,
                                                  const SState* __restrict__ ST, int* SIn,
                                                  int* bcnt, int* BI, int* DI, int* DIb,
                                                  LevelCounters* cnt) {
    const int nT_ = pcnt->nT;
    const int nblk_ = nT_;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_s_collect_one(bx, tasks, pcnt, first, SIp, T, ST, SIn, bcnt, BI, DI, DIb, cnt);
    }
}

// Final vectors: Z(isupp) scaled by 1/||z|| (the last dlar1v's NRMINV), zeros in the
// rest of the column (process_singleton.c:300-324).  One wave per task of the first
// iteration, which lists every singleton of the level.
__device__ __forceinline__ void k_s_normalize_one(int bx, const WaveTask* __restrict__ tasks,
                                                    const LevelCounters* __restrict__ cnt,
                                                    const STask* __restrict__ T,
                                                    const MatDesc* __restrict__ md,
                                                    MatState* __restrict__ ms,
                                                    const int* __restrict__ Zidx,
                                                    const SState* __restrict__ ST) {
    __shared__ double* zc_s[SW];
    __shared__ int lo_s[SW], hi_s[SW];
    __shared__ double sc_s[SW];
    if ((int)bx >= cnt->nT) return;
    const WaveTask wt = tasks[bx];
    const int lane = threadIdx.x;
    const int cntw = wt.cnt;
    const int idx = wt.start + min(lane, cntw - 1);
    const STask t = T[idx];
    const MatDesc& dsc = md[t.mat];
    const int col = Zidx[t.slot];
    if (__any(col >= dsc.mcols)) {
        // never past the caller's Z (k_zindex rules it out): write nothing, fail
        ms[t.mat].status = kTooFew;
        return;
    }
    const SState q = ST[idx];
    zc_s[lane] = dsc.Z + (size_t)col * dsc.ldz;
    lo_s[lane] = t.bbeg + q.is0 - 1;
    hi_s[lane] = t.bbeg + q.is1 - 1;
    sc_s[lane] = q.nrminv;
    const int nfull = dsc.n;
    __syncthreads();
    // four rows per lane in flight; only the support is read back
    for (int c = 0; c < cntw; ++c) {
        const GPtrW zc = as_global_w(zc_s[c]);
        const int lo = lo_s[c], hi = hi_s[c];
        const double sc = sc_s[c];
        for (int base = 0; base < nfull; base += 4 * SW) {
            double v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int rr = base + u * SW + lane;
                v[u] = (rr < nfull && rr >= lo && rr <= hi) ? zc[rr] * sc : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int rr = base + u * SW + lane;
                if (rr < nfull) zc[rr] = v[u];
            }
        }
    }
}

__global__ __launch_bounds__(64) void k_s_normalize(const WaveTask* __restrict__ tasks,
                                                    const LevelCounters* __restrict__ cnt,
                                                    const STask* __restrict__ T,
                                                    const MatDesc* __restrict__ md,
                                                    MatState* __restrict__ ms,
                                                    const int* __restrict__ Zidx,
                                                    const SState* __restrict__ ST) {
    const int nT_ = cnt->nT;
    const int nblk_ = nT_;
    for (int bx = blockIdx.x; bx < nblk_; bx += gridDim.x) {
        k_s_normalize_one(bx, tasks, cnt, T, md, ms, Zidx, ST);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// k_finalize: unscale (mrrr.c:371-386), place eigenvalues in column order
// (sort_eigenpairs, mrrr.c:525-588) and report the outputs.  One workgroup per
// matrix; the ordering check that may swap columns runs on one lane.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finalize(const MatDesc* __restrict__ md, MatState* ms,
                                                  const double* __restrict__ W,
                                                  const int* __restrict__ Zidx,
                                                  const BlockRec* __restrict__ blk,
                                                  const int* __restrict__ iblock,
                                                  const int* __restrict__ Windex, int* perm_all,
                                                  int* mark_all, double* colbuf_all) {
    __shared__ int sh_changed, sh_unsorted;
    __shared__ int sh_cand[256];
    __shared__ int sh_ncand;
    const int k = blockIdx.x;
    const int tid = threadIdx.x;
    MatState& S = ms[k];
    const MatDesc dsc = md[k];
    if (S.done || S.status != kOk) return;
    const int off = dsc.off, m = S.m;
    const double inv = 1.0 / S.scale;
    const bool wantz = dsc.jobz == kV;
    if (wantz && m > dsc.mcols) {
        // more eigenvectors than the caller's Z holds (k_zindex has failed such a
        // problem already): nothing is moved, so that no access reaches past Z or Zsupp
        if (tid == 0) { S.status = kTooFew; S.done = 1; }
        return;
    }
    const GPtrW Wout = as_global_w(dsc.Wout);
    for (int s = tid; s < m; s += blockDim.x) {
        double w = W[off + s];
        if (S.scale != 1.0) w *= inv;
        int pos = s;
        if (wantz) pos = Zidx[off + s];
        else if (S.nsplit > 1) {
            // jobz = N: rank in the merged ascending order (qsort with cmpa, mrrr.c:287)
            int rank = 0;
            for (int t = 0; t < m; ++t) {
                double v = W[off + t];
                if (v < W[off + s] || (v == W[off + s] && t < s)) ++rank;
            }
            pos = rank;
        }
        Wout[pos] = w;
    }
    if (tid == 0) sh_unsorted = 0;
    __syncthreads();
    if (tid == 0 && S.scale != 1.0) { S.vl *= inv; S.vu *= inv; }
    if (wantz) {
        // The ordering check of sort_eigenpairs (mrrr.c:557-586): passes that swap
        // neighbours with W(i) > W(i+1) until none does, moving Zsupp and the Z columns
        // along.  That is the stable sort by W of the order Zindex gives, which the
        // odd-even transposition sort below reaches too (it also swaps only strictly
        // larger neighbours); the columns then move once, along the cycles of the
        // resulting permutation.  Almost always nothing is out of order.
        for (int i = tid; i < m - 1; i += blockDim.x)
            if (Wout[i] > Wout[i + 1]) sh_unsorted = 1;
        __syncthreads();
        if (sh_unsorted) {
            int* perm = perm_all + off;    // perm[q]: the column that ends at position q
            int* mark = mark_all + off;
            const GPtrW colbuf = as_global_w(colbuf_all + off);
            const GIntW Zs = as_global_w(dsc.Zsupp);
            for (int q = tid; q < m; q += blockDim.x) { perm[q] = q; mark[q] = 0; }
            __syncthreads();
            int prev = 1;
            for (int phase = 0;; ++phase) {
                if (tid == 0) sh_changed = 0;
                __syncthreads();
                for (int i = 2 * tid + (phase & 1); i < m - 1; i += 2 * blockDim.x) {
                    const double a = Wout[i], b = Wout[i + 1];
                    if (a > b) {
                        Wout[i] = b;
                        Wout[i + 1] = a;
                        const int s0 = Zs[2 * i], s1 = Zs[2 * i + 1];
                        Zs[2 * i] = Zs[2 * i + 2];
                        Zs[2 * i + 1] = Zs[2 * i + 3];
                        Zs[2 * i + 2] = s0;
                        Zs[2 * i + 3] = s1;
                        const int p = perm[i];
                        perm[i] = perm[i + 1];
                        perm[i + 1] = p;
                        sh_changed = 1;
                    }
                }
                __syncthreads();
                const int ch = sh_changed;
                __syncthreads();
                if (!ch && !prev) break;   // an even and an odd phase without a swap
                prev = ch;
            }
            // move the columns along the permutation's cycles, one column of temporary
            const GPtrW Z = as_global_w(dsc.Z);
            const size_t ldz = (size_t)dsc.ldz;
            const int n = dsc.n;
            for (int base = 0; base < m; base += 256) {
                if (tid == 0) sh_ncand = 0;
                __syncthreads();
                const int q = base + tid;
                // candidates in index order: every thread's flag, then one thread lists them
                sh_cand[tid] = (q < m && perm[q] != q) ? 1 : 0;
                __syncthreads();
                if (tid == 0) {
                    int c = 0;
                    for (int t = 0; t < 256; ++t)
                        if (sh_cand[t]) sh_cand[c++] = base + t;
                    sh_ncand = c;
                }
                __syncthreads();
                const int nc = sh_ncand;
                for (int ci = 0; ci < nc; ++ci) {
                    const int start = sh_cand[ci];
                    if (mark[start]) continue;   // uniform: every thread reads the same flag
                    for (int r = tid; r < n; r += blockDim.x) colbuf[r] = Z[(size_t)start * ldz + r];
                    __syncthreads();
                    int cur = start;
                    for (;;) {
                        const int src = perm[cur];
                        if (tid == 0) mark[cur] = 1;
                        if (src == start) {
                            for (int r = tid; r < n; r += blockDim.x) Z[(size_t)cur * ldz + r] = colbuf[r];
                            __syncthreads();
                            break;
                        }
                        for (int r = tid; r < n; r += blockDim.x)
                            Z[(size_t)cur * ldz + r] = Z[(size_t)src * ldz + r];
                        __syncthreads();
                        cur = src;
                    }
                }
                __syncthreads();
            }
        }
    }
    if (tid == 0) S.done = 1;
}

}  // namespace mr3
