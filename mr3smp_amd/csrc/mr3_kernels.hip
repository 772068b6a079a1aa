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
//   process_singleton.c:65-324 S-task / dlar1v         -> k_singleton (one lane per vector)
//   mrrr.c:263-288 shift back, dlarrj, unscale, sort   -> k_shift_n, k_rrj, k_finalize
#include <hip/hip_runtime.h>

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
    kStepKinds = 10
};
__device__ unsigned long long g_steps[kStepKinds];
__constant__ int g_count_steps;

__device__ __forceinline__ void add_steps(int kind, unsigned long long v) {
    if (g_count_steps && v) atomicAdd(&g_steps[kind], v);
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

// dlaebz IJOB=3 for one of the two target counts (mrrr_val.c:295-319); the two
// intervals of the reference never interact, so each runs on its own lane.
__device__ void ijob3_one(int n, const double* d, const double* e2, double pivmin,
                          double abstol, double reltol, int nitmax, int nval, double c,
                          double* a1, double* a2, int* n1, int* n2) {
    for (int it = 0; it < nitmax; ++it) {
        int cnt = count_T_ijob2(n, d, e2, c, pivmin);
        if (cnt <= nval) { *a1 = c; *n1 = cnt; }
        if (cnt >= nval) { *a2 = c; *n2 = cnt; }
        double tmp1 = fabs(*a2 - *a1), tmp2 = fmax(fabs(*a2), fabs(*a1));
        if (tmp1 < fmax(fmax(abstol, pivmin), reltol * tmp2) || *n1 >= *n2) break;
        c = 0.5 * (*a1 + *a2);
    }
}

// ---------------------------------------------------------------------------------
// k_prep: scaling, rel-acc test, Gershgorin, pivmin, splitting, counts and the
// per-block index ranges of one matrix (mrrr.c:155-243, mrrr_val.c:110-438).
// ---------------------------------------------------------------------------------
constexpr int PREP_NT = 256;

__global__ __launch_bounds__(PREP_NT) void k_prep(const MatDesc* __restrict__ md,
                                                  MatState* __restrict__ ms, double* D, double* E,
                                                  double* E2, double* gers, double* Dcopy,
                                                  double* E2copy, int* isplit, BlockRec* blk,
                                                  double* W, double* Werr, double* Wcoarse,
                                                  int* iblock, int* Windex, int* smat) {
    __shared__ double shd[PREP_NT];
    __shared__ int shi[PREP_NT];
    __shared__ double sh_wl, sh_wu, sh_wlu, sh_wul;
    __shared__ int sh_nwl, sh_nwu, sh_c0, sh_c1;
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
        if (tid < 2) {
            double a1 = gl, a2 = gu;
            int n1 = -1, n2 = n + 1;
            ijob3_one(n, d, e2, pivmin, atoli, rqtol, itmax, tid == 0 ? il - 1 : iu,
                      tid == 0 ? gl : gu, &a1, &a2, &n1, &n2);
            if (tid == 0) { sh_wl = a1; sh_wlu = a2; sh_nwl = n1; }
            else { sh_wu = a2; sh_wul = a1; sh_nwu = n2; }
        }
        __syncthreads();
        wl = sh_wl; wlu = sh_wlu; nwl = sh_nwl;
        wu = sh_wu; wul = sh_wul; nwu = sh_nwu;
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
// k_coarse: eigenvalue approximations by bisection on T, one lane per wanted
// eigenvalue (dlarrd/dlaebz, odrrd.f:415-574, to RELTOL = bsrtol).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_coarse(const MatDesc* __restrict__ md,
                                                const MatState* __restrict__ ms,
                                                const double* __restrict__ D,
                                                const double* __restrict__ E2,
                                                const BlockRec* __restrict__ blk,
                                                const int* __restrict__ smat,
                                                const int* __restrict__ iblock,
                                                const int* __restrict__ Windex, double* W,
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
    // dlarrd called on the block alone (mrrr_val.c:1291-1295): its own fudged
    // bounds (odrrd.f:515-519 with FUDGE = 2) and tolerances (odrrd.f:337-341).
    const double tnormb = fmax(fabs(br.glraw), fabs(br.guraw));
    double left = br.glraw - 2.0 * tnormb * kEps * nbl - 2.0 * pivmin;
    double right = br.guraw + 2.0 * tnormb * kEps * nbl + 2.0 * pivmin;
    const double atoli = 2.0 * 2.0 * kSafmin + 2.0 * 2.0 * pivmin;
    const double rtoli = S.bsrtol;
    const int itmax = (int)((log(right - left + pivmin) - log(pivmin)) / log(2.0)) + 2;
    int nl = 0, nr = nbl, sweeps = 0;
    for (int it = 0; it < itmax; ++it) {
        double tmp1 = fabs(right - left), tmp2 = fmax(fabs(right), fabs(left));
        if (tmp1 < fmax(fmax(atoli, pivmin), rtoli * tmp2)) break;
        double mid = 0.5 * (left + right);
        int c = count_T_ijob2(nbl, d, e2, mid, pivmin);
        ++sweeps;
        c = min(nr, max(nl, c));
        if (c >= i) { right = mid; nr = c; }
        else { left = mid; nl = c; }
    }
    add_steps(kStepCoarse, (unsigned long long)sweeps * nbl);
    const double w = 0.5 * (left + right);
    W[gidx] = w;
    Wcoarse[gidx] = w;
    Werr[gidx] = 0.5 * fabs(left - right);
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
__global__ __launch_bounds__(64) void k_root(const MatDesc* __restrict__ md, MatState* ms,
                                             BlockRec* blk, const int* __restrict__ smat,
                                             const int* __restrict__ blist, double* D, double* E,
                                             double* DLr, double* DLLr, double* W, double* Werr,
                                             double* Wgap, int nblist) {
    if ((int)blockIdx.x >= nblist) return;
    const int gb = blist[blockIdx.x];
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
    __shared__ double sh_sigma, sh_spdiam;
    __shared__ int sh_fail;
    // gaps of the coarse values (mrrr_val.c:431-435)
    for (int s = wbgn + lane; s <= wend; s += 64) {
        if (s < wend) Wgap[s] = fmax(0.0, (W[s + 1] - Werr[s + 1]) - (W[s] + Werr[s]));
        else Wgap[s] = fmax(0.0, S.wu - (W[s] + Werr[s]));
    }
    __syncthreads();
    if (lane == 0) {
        const double gl = br.gl, gu = br.gu, vl = S.vl, vu = S.vu;
        const int ilow = br.ilow, iupp = br.iupp;
        double isleft = fmax(gl, (W[wbgn] - Werr[wbgn]) - 100.0 * kEps * fabs(W[wbgn] - Werr[wbgn]));
        double isright = fmin(gu, (W[wend] + Werr[wend]) + 100.0 * kEps * fabs(W[wend] - Werr[wend]));
        double spdiam = gu - gl;
        double tmp0 = fmin(isright, vu) - fmax(isleft, vl);
        double s1 = fmax(isleft, vl) + 0.25 * tmp0;
        double s2 = fmin(isright, vu) - 0.25 * tmp0;
        int lcnt = 0, rcnt = 0;
        if (mbl > 1) { // dlarrc 'T' (odrrc.f:95-113)
            double lp = d[0] - s1, rp = d[0] - s2;
            lcnt = (lp <= 0.0); rcnt = (rp <= 0.0);
            for (int i = 0; i < nbl - 1; ++i) {
                double t = e[i] * e[i];
                lp = (d[i + 1] - s1) - t / lp;
                rp = (d[i + 1] - s2) - t / rp;
                lcnt += (lp <= 0.0); rcnt += (rp <= 0.0);
            }
        }
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
        int ok = 0;
        for (int jtry = 0; jtry < kMaxTryRoot; ++jtry) {
            double dpivot = d[0] - sigma;
            dl[0] = dpivot;                 // D of the trial factorisation
            double dmax = fabs(dpivot);
            for (int i = 0; i < nbl - 1; ++i) {
                double invd = 1.0 / dl[i];
                double t = e[i] * invd;
                dll[i] = t;                 // L of the trial factorisation
                dpivot = (d[i + 1] - sigma) - t * e[i];
                dl[i + 1] = dpivot;
                dmax = fmax(dmax, fabs(dpivot));
            }
            if (dmax > kMaxGrowth * spdiam) {
                if (jtry == kMaxTryRoot - 2) {
                    if (sgndef == 1) sigma = gl - kFudgeRoot * spdiam * kEps * nbl - kFudgeRoot * 2.0 * pivmin;
                    else sigma = gu + kFudgeRoot * spdiam * kEps * nbl + kFudgeRoot * 2.0 * pivmin;
                } else if (jtry == kMaxTryRoot - 1) {
                    break;
                } else {
                    sigma -= sgndef * tau;
                    tau *= 2.0;
                }
            } else { ok = 1; break; }
        }
        sh_fail = !ok;
        sh_sigma = sigma;
        sh_spdiam = spdiam;
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
        int sw = dlarrb1(nbl, D + off + br.bbeg, DLLr + off + br.bbeg, Windex[gidx], S.rtol1,
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
    int sw = dlarrj1(nbl, Dcopy + off + br.bbeg, E2copy + off + br.bbeg, Windex[gidx],
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
                                                const MatState* __restrict__ ms,
                                                const BlockRec* __restrict__ blk,
                                                const double* __restrict__ W,
                                                const int* __restrict__ iblock,
                                                const int* __restrict__ Windex, int* Zidx) {
    const int k = blockIdx.x;
    const MatState& S = ms[k];
    if (S.done || S.status != kOk || md[k].jobz != kV) return;
    const int off = md[k].off, m = S.m, nsplit = S.nsplit;
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
    int capS, capC, capEig;
};

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
            t.bl_spdiam = bl_spdiam;
            t.lgap = (new_first == cb) ? cl_lgap : Wgap[new_first - 1];
            t.D = rD; t.L = rL; t.DL = rDL; t.DLL = rDLL;
            out.S[sbase++] = t;
        } else {
            auto emit = [&](int f, int l) {
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
                for (int j = f; j <= l; ++j) {
                    out.eig_slot[ebase] = j;
                    out.eig_cl[ebase] = cbase;
                    ++ebase;
                }
                out.C[cbase++] = c;
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

// Level 0: one thread per block slot.  1x1 blocks become singletons of size 1
// (mrrr_vec.c:330-346).
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
                                                      TreeOut out, int nblist) {
    const int ib = blockIdx.x * blockDim.x + threadIdx.x;
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
    if (nbl == 1) {
        atomicAdd(&out.cnt->nS, 1);
        int pos = off + br.wbeg + atomicAdd(&out.bcnt[gb], 1);
        atomicMax(&out.cnt->maxbl, 1);
        if (pos >= out.capS) { atomicOr(&out.cnt->err, 1); return; }
        STask t;
        t.slot = cb; t.mat = k; t.bbeg = br.bbeg; t.bsize = 1; t.bl_spdiam = 0.0; t.lgap = 0.0;
        t.D = rD; t.L = rL; t.DL = DLr + off + br.bbeg; t.DLL = DLLr + off + br.bbeg;
        out.S[pos] = t;
        Wsh[cb] = W[cb];
        return;
    }
    // refine_eigvals at depth 0 (process_cluster.c:265-281): the spread of the block
    // from gersch rows bl_begin .. bl_end-1, Wshifted = W, W += sigma.
    const double* g = gers + 2 * (size_t)off;
    double gl = g[2 * br.bbeg], gu = g[2 * br.bbeg + 1];
    for (int i = br.bbeg + 1; i < br.bend; ++i) { gl = fmin(gl, g[2 * i]); gu = fmax(gu, g[2 * i + 1]); }
    const double bl_spdiam = gu - gl;
    const double sigma = br.sigma;
    for (int s = cb; s <= ce; ++s) { Wsh[s] = W[s]; W[s] += sigma; }
    // lgap of the root task (mrrr_vec.c:357)
    const double lgap = fmax(0.0, (Wsh[cb] + sigma) - Werr[cb] - S.vl);
    classify_cluster(cb, ce, 0, k, gb, off + br.wbeg, br.bbeg, nbl, bl_spdiam, lgap, rD, rL,
                     DLr + off + br.bbeg, DLLr + off + br.bbeg, Wgap, Wsh, out);
}

// Deeper levels: one thread per cluster, after its eigenvalues were refined.
__global__ __launch_bounds__(64) void k_classify(const ClTask* __restrict__ C, int nC,
                                                 const double* __restrict__ Wgap,
                                                 const double* __restrict__ Wsh, TreeOut out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    const ClTask t = C[c];
    classify_cluster(t.cb, t.ce, t.depth, t.mat, t.gblk, t.wseg, t.bbeg, t.bsize, t.bl_spdiam,
                     t.lgap, t.cD, t.cL, t.cDL, t.cDLL, Wgap, Wsh, out);
}

// Cuts every block's singletons of this level into wave tasks of at most 64 and
// clears the block counters for the next level.  One thread per block.
__global__ __launch_bounds__(64) void k_plan_s(const int* __restrict__ blist, int nblist,
                                               const int* __restrict__ smat,
                                               const MatDesc* __restrict__ md,
                                               const BlockRec* __restrict__ blk, int* bcnt,
                                               WaveTask* tasks, int capT, LevelCounters* cnt) {
    const int ib = blockIdx.x * blockDim.x + threadIdx.x;
    if (ib >= nblist) return;
    const int gb = blist[ib];
    const int c = bcnt[gb];
    if (c == 0) return;
    bcnt[gb] = 0;
    const int wseg = md[smat[gb]].off + blk[gb].wbeg;
    const int k = (c + 63) / 64;
    const int base = atomicAdd(&cnt->nT, k);
    if (base + k > capT) { atomicOr(&cnt->err, 2); return; }
    for (int t = 0; t < k; ++t) tasks[base + t] = WaveTask{wseg + 64 * t, min(64, c - 64 * t)};
}

// Assigns each cluster of a level its arena slot (4 * bsize doubles).
__global__ __launch_bounds__(256) void k_arena_assign(ClTask* C, int nC, double* arena,
                                                      unsigned long long* bump) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    ClTask& t = C[c];
    unsigned long long o = atomicAdd(bump, (unsigned long long)(4LL * t.bsize));
    double* base = arena + o;
    t.arena_off = (long long)o;
    t.cD = base;
    t.cL = base + t.bsize;
    t.cDL = base + 2 * t.bsize;
    t.cDLL = base + 3 * t.bsize;
}

// ---------------------------------------------------------------------------------
// k_newrrr: compute_new_rrr (process_cluster.c:105-235) for one cluster at depth >= 1.
// Lanes 0 and 1 refine the two extremal eigenvalues against the parent (the two
// dlarrb calls are independent: each runs with a zero gap); lane 0 then runs dlarrf
// (odrrf.f:140-384); all lanes apply the shift and form DL, DLL.
// ---------------------------------------------------------------------------------
__device__ int dlarrf_lane(int n, const double* __restrict__ d, const double* __restrict__ l,
                           const double* __restrict__ ld, const double* w, const double* wgap,
                           const double* werr, int clsize, double spdiam, double clgapl,
                           double clgapr, double pivmin, double* sigma, double* dplus,
                           double* lplus, double* work) {
    // w/wgap/werr point at the cluster's first eigenvalue; CLSTRT = 1, CLEND = clsize.
    const int KTRYMAX = 1;
    const double fact = (double)(1 << KTRYMAX);
    int shift = 0, forcer = 0, indx = 0;
    const double clwdth = fabs(w[clsize - 1] - w[0]) + werr[clsize - 1] + werr[0];
    const double avgap = clwdth / (double)(clsize - 1);
    const double mingap = fmin(clgapl, clgapr);
    double lsigma = fmin(w[0], w[clsize - 1]) - werr[0];
    double rsigma = fmax(w[0], w[clsize - 1]) + werr[clsize - 1];
    lsigma = lsigma - fabs(lsigma) * 4.0 * kEps;
    rsigma = rsigma + fabs(rsigma) * 4.0 * kEps;
    const double ldmax = 0.25 * mingap + 2.0 * pivmin;
    const double rdmax = 0.25 * mingap + 2.0 * pivmin;
    double ldelta = fmax(avgap, wgap[0]) / fact;
    double rdelta = fmax(avgap, wgap[clsize - 2]) / fact;
    double smlgrowth = 1.0 / kSafmin;
    const double fail = (double)(n - 1) * mingap / (spdiam * kEps);
    const double fail2 = (double)(n - 1) * mingap / (spdiam * sqrt(kEps));
    double bestshift = lsigma;
    int ktry = 0;
    const double growthbound = kMaxGrowth1 * spdiam;
    double max1 = 0.0, max2 = 0.0;
    int sweeps = 0;
    for (;;) {
        ++sweeps;
        int sawnan1 = 0, sawnan2 = 0;
        ldelta = fmin(ldmax, ldelta);
        rdelta = fmin(rdmax, rdelta);
        double s = -lsigma;
        dplus[0] = d[0] + s;
        if (fabs(dplus[0]) < pivmin) { dplus[0] = -pivmin; sawnan1 = 1; }
        max1 = fabs(dplus[0]);
        for (int i = 0; i < n - 1; ++i) {
            lplus[i] = ld[i] / dplus[i];
            s = s * lplus[i] * l[i] - lsigma;
            dplus[i + 1] = d[i + 1] + s;
            if (fabs(dplus[i + 1]) < pivmin) { dplus[i + 1] = -pivmin; sawnan1 = 1; }
            max1 = fmax(max1, fabs(dplus[i + 1]));
        }
        sawnan1 = sawnan1 || isnan(max1);
        if (forcer || (max1 <= growthbound && !sawnan1)) { *sigma = lsigma; shift = 1; break; }
        ++sweeps;
        s = -rsigma;
        work[0] = d[0] + s;
        if (fabs(work[0]) < pivmin) { work[0] = -pivmin; sawnan2 = 1; }
        max2 = fabs(work[0]);
        for (int i = 0; i < n - 1; ++i) {
            work[n + i] = ld[i] / work[i];
            s = s * work[n + i] * l[i] - rsigma;
            work[i + 1] = d[i + 1] + s;
            if (fabs(work[i + 1]) < pivmin) { work[i + 1] = -pivmin; sawnan2 = 1; }
            max2 = fmax(max2, fabs(work[i + 1]));
        }
        sawnan2 = sawnan2 || isnan(max2);
        if (forcer || (max2 <= growthbound && !sawnan2)) { *sigma = rsigma; shift = 2; break; }
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
                if (indx == 1) {
                    double tmp = fabs(dplus[n - 1]), znm2 = 1.0, prod = 1.0, oldp = 1.0;
                    for (int i = n - 2; i >= 0; --i) {
                        if (prod <= kEps) prod = ((dplus[i + 1] * work[n + i + 1]) / (dplus[i] * work[n + i])) * oldp;
                        else prod = prod * fabs(work[n + i]);
                        oldp = prod;
                        znm2 = znm2 + prod * prod;
                        tmp = fmax(tmp, fabs(dplus[i] * prod));
                    }
                    double rrr1 = tmp / (spdiam * sqrt(znm2));
                    if (rrr1 <= kMaxGrowth2) { *sigma = lsigma; shift = 1; break; }
                } else if (indx == 2) {
                    double tmp = fabs(work[n - 1]), znm2 = 1.0, prod = 1.0, oldp = 1.0;
                    for (int i = n - 2; i >= 0; --i) {
                        if (prod <= kEps) prod = ((work[i + 1] * lplus[i + 1]) / (work[i] * lplus[i])) * oldp;
                        else prod = prod * fabs(lplus[i]);
                        oldp = prod;
                        znm2 = znm2 + prod * prod;
                        tmp = fmax(tmp, fabs(work[i] * prod));
                    }
                    double rrr2 = tmp / (spdiam * sqrt(znm2));
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
        (void)fail;
    }
    add_steps(kStepRrf, (unsigned long long)sweeps * n);
    return shift;
}

// Extremal eigenvalues of every cluster to full accuracy against the parent
// (process_cluster.c:168-195), one lane per cluster end.  The two dlarrb calls of a
// cluster run with a zero gap and do not depend on each other; results go to Wtmp
// and are committed by k_dlarrf.
__global__ __launch_bounds__(64) void k_endref(const ClTask* __restrict__ C, int nC,
                                               const int* __restrict__ Windex,
                                               const double* __restrict__ Wsh,
                                               const double* __restrict__ Werr,
                                               const MatState* __restrict__ ms, double* Wtmp,
                                               double* Werrtmp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 2 * nC) return;
    const ClTask t = C[e >> 1];
    const int slot = (e & 1) ? t.ce : t.cb;
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    int sw = dlarrb1(t.bsize, t.pD, t.pDLL, Windex[slot], S.rqtol, S.rqtol, 0.0, &w, &we,
                     S.pivmin, t.bl_spdiam, t.bsize);
    add_steps(kStepEndref, (unsigned long long)sw * t.bsize);
    Wtmp[slot] = w;
    Werrtmp[slot] = we;
}

// New representation of every cluster by dlarrf (odrrf.f:140-384), one lane per
// cluster.  The left-shift factorisation goes to cD/cL, the right-shift one to the
// cDL/cDLL space; k_newrrr_fin keeps the chosen one.
__global__ __launch_bounds__(64) void k_dlarrf(ClTask* C, int nC, double* Wsh, double* Werr,
                                               double* Wgap, const double* __restrict__ Wtmp,
                                               const double* __restrict__ Werrtmp,
                                               const MatState* __restrict__ ms) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nC) return;
    ClTask& t = C[c];
    const int cb = t.cb, ce = t.ce, clsize = ce - cb + 1;
    // Wgap[cb] from the refined left end and the neighbour as it was before the right
    // end was refined (process_cluster.c:186-189), then commit both ends
    Wgap[cb] = fmax(0.0, (Wsh[cb + 1] - Werr[cb + 1]) - (Wtmp[cb] + Werrtmp[cb]));
    Wsh[cb] = Wtmp[cb];
    Werr[cb] = Werrtmp[cb];
    Wsh[ce] = Wtmp[ce];
    Werr[ce] = Werrtmp[ce];
    double tau;
    int sh = dlarrf_lane(t.bsize, t.pD, t.pL, t.pDL, Wsh + cb, Wgap + cb, Werr + cb, clsize,
                         t.bl_spdiam, t.lgap, Wgap[ce], ms[t.mat].pivmin, &tau, t.cD, t.cL,
                         t.cDL);
    t.tau = tau;
    t.shift = sh;
}

// Keeps the chosen factorisation, accumulates the shift, forms DL and DLL and moves
// the cluster's eigenvalues to the new representation (process_cluster.c:206-232).
__global__ __launch_bounds__(64) void k_newrrr_fin(const ClTask* __restrict__ C, int nC,
                                                   double* Wsh, double* Werr) {
    const int c = blockIdx.x;
    if (c >= nC) return;
    const ClTask t = C[c];
    const int n = t.bsize;
    const int lane = threadIdx.x;
    const bool right = (t.shift == 2);
    for (int i = lane; i < n; i += 64) {
        double dv, lv;
        if (right) {
            dv = t.cDL[i];                               // WORK(i+1)
            lv = (i < n - 1) ? t.cDL[n + i] : 0.0;       // WORK(N+i+1)
            t.cD[i] = dv;
            if (i < n - 1) t.cL[i] = lv;
        } else {
            dv = t.cD[i];
            lv = (i < n - 1) ? t.cL[i] : 0.0;
        }
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

// Refinement of a cluster's eigenvalues against its new representation, one lane
// per eigenvalue (process_cluster.c:414-418; the cluster is one dlarrb call).
__global__ __launch_bounds__(256) void k_refine_cl(const ClTask* __restrict__ C,
                                                   const int* __restrict__ eig_slot,
                                                   const int* __restrict__ eig_cl, int nE,
                                                   const int* __restrict__ Windex,
                                                   const double* __restrict__ Wsh,
                                                   const double* __restrict__ Werr,
                                                   const double* __restrict__ Wgap,
                                                   const MatState* __restrict__ ms, double* Wtmp,
                                                   double* Werrtmp) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int slot = eig_slot[e];
    const ClTask t = C[eig_cl[e]];
    const MatState& S = ms[t.mat];
    double w = Wsh[slot], we = Werr[slot];
    double rgap = Wgap[slot];
    double lgap = (slot == t.cb) ? rgap : Wgap[slot - 1];
    int sw = dlarrb1(t.bsize, t.cD, t.cDLL, Windex[slot], S.rtol1, S.rtol2, fmin(lgap, rgap),
                     &w, &we, S.pivmin, t.bl_spdiam, t.bsize);
    add_steps(kStepRefineCl, (unsigned long long)sw * t.bsize);
    Wtmp[slot] = w;
    Werrtmp[slot] = we;
}

__global__ __launch_bounds__(256) void k_gaps_cl(const ClTask* __restrict__ C,
                                                 const int* __restrict__ eig_slot,
                                                 const int* __restrict__ eig_cl, int nE,
                                                 const double* __restrict__ Wtmp,
                                                 const double* __restrict__ Werrtmp, double* Wsh,
                                                 double* Werr, double* Wgap, double* W) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int slot = eig_slot[e];
    const ClTask t = C[eig_cl[e]];
    const double w = Wtmp[slot], we = Werrtmp[slot];
    Wsh[slot] = w;
    Werr[slot] = we;
    if (slot < t.ce) Wgap[slot] = fmax(0.0, Wtmp[slot + 1] - Werrtmp[slot + 1] - w - we);
    W[slot] = w + t.cL[t.bsize - 1];
}

// ---------------------------------------------------------------------------------
// Eigenvectors: dlar1v and its vector solve for a wave of singletons of one block,
// then the singleton kernel (process_singleton.c:65-324; dlar1v odr1v.f:156-369).
// ---------------------------------------------------------------------------------
constexpr int SW = 64;          // lanes per wave
constexpr int TILE_R = 16;      // rows per LDS tile
constexpr int kUS = 4;          // chunk of the eigenvector sweeps (register budget)

__device__ __forceinline__ int wave_max_u(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}
__device__ __forceinline__ int wave_min_u(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return __builtin_amdgcn_readfirstlane(v);
}

// Lane-interleaved scratch of one wave with the three arrays interleaved per element:
// element i of LPLUS, UMINUS and S/gamma of lane l sit at ((3i + a) * 64 + l).  With
// i wave-uniform every access is one 512-byte row at a scalar base plus the lane.
struct WScr {
    double* base;
    int lane;
    __device__ __forceinline__ double& LP(int i) const { return base[(3 * i) * SW + lane]; }
    __device__ __forceinline__ double& UM(int i) const { return base[(3 * i + 1) * SW + lane]; }
    __device__ __forceinline__ double& SG(int i) const { return base[(3 * i + 2) * SW + lane]; }
};

struct R1v {
    int negcnt, r, isupp0, isupp1, sawnan;
    double ztz, mingma, nrminv, resid, rqcorr;
};

struct Rrr {
    const double* D;
    const double* L;
    const double* LD;
    const double* LLD;
};

__device__ void r1v_wave(int n, double lambda, const Rrr& R, double pivmin, int wantnc, int rin,
                         bool act, const WScr& sc, R1v* o) {
    const int r1 = rin ? rin : 1, r2 = rin ? rin : n;
    // stationary transform (odr1v.f:185-229); loops 50 and 51 back to back
    int neg1 = 0;
    double s = 0.0 - lambda;
    if (act) sc.SG(0) = 0.0;
    const int stop = act ? r2 : 1;            // this lane's steps: i = 1 .. stop-1
    const int stopU = wave_max_u(stop);
    if (stopU > 1) {
        double dd[kUS], ld[kUS], ll[kUS], dn[kUS], ldn[kUS], lln[kUS];
#pragma unroll
        for (int u = 0; u < kUS; ++u) {
            int j = min(1 + u, stopU - 1) - 1;
            dd[u] = R.D[j]; ld[u] = R.LD[j]; ll[u] = R.L[j];
        }
        for (int i0 = 1; i0 < stopU; i0 += kUS) {
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                int j = min(i0 + kUS + u, stopU - 1) - 1;
                dn[u] = R.D[j]; ldn[u] = R.LD[j]; lln[u] = R.L[j];
            }
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                const int i = i0 + u;
                if (i < stop) {
                    double dplus = dd[u] + s;
                    double lp = ld[u] / dplus;
                    sc.LP(i) = lp;
                    neg1 += (i <= r1 - 1) && (dplus < 0.0);
                    double si = s * lp * ll[u];
                    sc.SG(i) = si;
                    s = si - lambda;
                }
            }
#pragma unroll
            for (int u = 0; u < kUS; ++u) { dd[u] = dn[u]; ld[u] = ldn[u]; ll[u] = lln[u]; }
        }
    }
    const int sawnan1 = act && isnan(s);
    if (sawnan1) {
        neg1 = 0;
        s = 0.0 - lambda;
        for (int i = 1; i <= r2 - 1; ++i) {
            double dplus = R.D[i - 1] + s;
            if (fabs(dplus) < pivmin) dplus = -pivmin;
            double lp = R.LD[i - 1] / dplus;
            sc.LP(i) = lp;
            if (i <= r1 - 1) neg1 += (dplus < 0.0);
            double si = s * lp * R.L[i - 1];
            if (lp == 0.0) si = R.LLD[i - 1];
            sc.SG(i) = si;
            s = si - lambda;
        }
    }
    // progressive transform (odr1v.f:234-260)
    int neg2 = 0;
    const double pn = R.D[n - 1] - lambda;   // P(BN-1)
    double gam_r1m1 = 0.0;
    double pi = pn;
    if (act) {
        if (r1 - 1 == n - 1) gam_r1m1 = sc.SG(n - 1) + pi;
        else if (rin == 0) {
            double gm = sc.SG(n - 1) + pi;
            if (gm == 0.0) gm = kEps * sc.SG(n - 1);
            sc.SG(n - 1) = gm;
        }
    }
    const int lo = act ? r1 : n;               // this lane's steps: i = n-1 down to lo
    const int loU = wave_min_u(lo);
    if (n - 1 >= loU) {
        double dd[kUS], ll[kUS], lld[kUS], sv[kUS], dn[kUS], lln[kUS], lldn[kUS], svn[kUS];
#pragma unroll
        for (int u = 0; u < kUS; ++u) {
            int i = max(n - 1 - u, loU);
            dd[u] = R.D[i - 1]; ll[u] = R.L[i - 1]; lld[u] = R.LLD[i - 1]; sv[u] = sc.SG(i - 1);
        }
        for (int i0 = n - 1; i0 >= loU; i0 -= kUS) {
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                int i = max(i0 - kUS - u, loU);
                dn[u] = R.D[i - 1]; lln[u] = R.L[i - 1]; lldn[u] = R.LLD[i - 1]; svn[u] = sc.SG(i - 1);
            }
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                const int i = i0 - u;
                if (i >= lo) {
                    double dminus = lld[u] + pi;
                    double tmp = dd[u] / dminus;
                    neg2 += (dminus < 0.0);
                    sc.UM(i) = ll[u] * tmp;
                    double pim1 = pi * tmp - lambda;   // P(I-1)
                    if (i - 1 == r1 - 1) gam_r1m1 = sv[u] + pim1;
                    else if (rin == 0) {
                        double gm = sv[u] + pim1;
                        if (gm == 0.0) gm = kEps * sv[u];
                        sc.SG(i - 1) = gm;
                    }
                    pi = pim1;
                }
            }
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                dd[u] = dn[u]; ll[u] = lln[u]; lld[u] = lldn[u]; sv[u] = svn[u];
            }
        }
    }
    const int sawnan2 = act && isnan(pi);
    if (sawnan2) {
        // slow progressive form (odr1v.f:240-258); S is formed again because the fast
        // sweep overwrote it with gamma values
        neg2 = 0;
        double ss = 0.0 - lambda;
        sc.SG(0) = 0.0;
        for (int i = 1; i <= r2 - 1; ++i) {
            double dplus = R.D[i - 1] + ss;
            if (sawnan1 && fabs(dplus) < pivmin) dplus = -pivmin;
            double lp = R.LD[i - 1] / dplus;
            double si = ss * lp * R.L[i - 1];
            if (sawnan1 && lp == 0.0) si = R.LLD[i - 1];
            sc.SG(i) = si;
            ss = si - lambda;
        }
        pi = pn;
        if (r1 - 1 == n - 1) gam_r1m1 = sc.SG(n - 1) + pi;
        else if (rin == 0) {
            double gm = sc.SG(n - 1) + pi;
            if (gm == 0.0) gm = kEps * sc.SG(n - 1);
            sc.SG(n - 1) = gm;
        }
        for (int i = n - 1; i >= r1; --i) {
            double dminus = R.LLD[i - 1] + pi;
            if (fabs(dminus) < pivmin) dminus = -pivmin;
            double tmp = R.D[i - 1] / dminus;
            neg2 += (dminus < 0.0);
            sc.UM(i) = R.L[i - 1] * tmp;
            double pim1 = pi * tmp - lambda;
            if (tmp == 0.0) pim1 = R.D[i - 1] - lambda;
            if (i - 1 == r1 - 1) gam_r1m1 = sc.SG(i - 1) + pim1;
            else if (rin == 0) {
                double gm = sc.SG(i - 1) + pim1;
                if (gm == 0.0) gm = kEps * sc.SG(i - 1);
                sc.SG(i - 1) = gm;
            }
            pi = pim1;
        }
    }
    // twist index (odr1v.f:265-283): lanes with a free twist scan I = 1 .. n-1
    double mingma = gam_r1m1;
    if (act && mingma < 0.0) ++neg1;
    if (act && fabs(mingma) == 0.0) mingma = kEps * sc.SG(r1 - 1);
    int r = r1;
    const bool scan = act && rin == 0;
    if (__any(scan) && n > 1) {
        double g[kUS], gn[kUS];
#pragma unroll
        for (int u = 0; u < kUS; ++u) g[u] = sc.SG(min(1 + u, n - 1));
        for (int i0 = 1; i0 <= n - 1; i0 += kUS) {
#pragma unroll
            for (int u = 0; u < kUS; ++u) gn[u] = sc.SG(min(i0 + kUS + u, n - 1));
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                if (scan && i0 + u <= n - 1 && fabs(g[u]) <= fabs(mingma)) {
                    mingma = g[u];
                    r = i0 + u + 1;
                }
            }
#pragma unroll
            for (int u = 0; u < kUS; ++u) g[u] = gn[u];
        }
    }
    if (act) {
        o->negcnt = wantnc ? neg1 + neg2 : -1;
        o->r = r;
        o->mingma = mingma;
        o->sawnan = sawnan1 || sawnan2;
    }
}

// The vector solve of odr1v.f:287-355 for every lane with act set.  With stage ==
// false it forms ZTZ and the support; with stage == true it writes z * scale for the
// block rows 1..n into SG (zero outside the support).  Index ranges are wave-uniform
// as in r1v_wave.
__device__ void solve_wave(int n, const double* __restrict__ LD, double gaptol, const WScr& sc,
                           R1v* o, bool act, bool stage, double scale) {
    const int r = act ? o->r : 1;
    const bool nan = act && o->sawnan;
    int is0 = 1, is1 = n;
    double ztz = 1.0;
    if (act && stage) sc.SG(r - 1) = scale;   // Z(r) = 1
    // upwards from r: this lane covers I = r-1 down to 1
    {
        double zi1 = 1.0, zi2 = 0.0;   // Z(I+1), Z(I+2)
        bool live = act && !nan && r - 1 >= 1;
        const int topU = wave_max_u((act && !nan) ? r - 1 : 0);
        if (topU >= 1) {
            double lp[kUS], ld[kUS], lpn[kUS], ldn[kUS];
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                int i = max(topU - u, 1);
                lp[u] = sc.LP(i); ld[u] = LD[i - 1];
            }
            for (int i0 = topU; i0 >= 1; i0 -= kUS) {
                if (!stage && !__any(live)) break;
#pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    int i = max(i0 - kUS - u, 1);
                    lpn[u] = sc.LP(i); ldn[u] = LD[i - 1];
                }
#pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int i = i0 - u;
                    if (act && !nan && i >= 1 && i <= r - 1) {
                        if (live) {
                            double zi = -(lp[u] * zi1);
                            if ((fabs(zi) + fabs(zi1)) * fabs(ld[u]) < gaptol) {
                                is0 = i + 1;
                                live = false;
                                if (stage) sc.SG(i - 1) = 0.0;
                            } else {
                                ztz = ztz + zi * zi;
                                if (stage) sc.SG(i - 1) = zi * scale;
                                zi1 = zi;
                            }
                        } else if (stage) {
                            sc.SG(i - 1) = 0.0;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kUS; ++u) { lp[u] = lpn[u]; ld[u] = ldn[u]; }
            }
        }
        if (nan) {   // slow form after a NaN (odr1v.f:308-322)
            bool cut = false;
            for (int i = r - 1; i >= 1; --i) {
                if (cut) { if (stage) sc.SG(i - 1) = 0.0; continue; }
                double zi;
                if (zi1 == 0.0) zi = -(LD[i] / LD[i - 1]) * zi2;
                else zi = -(sc.LP(i) * zi1);
                if ((fabs(zi) + fabs(zi1)) * fabs(LD[i - 1]) < gaptol) {
                    is0 = i + 1;
                    cut = true;
                    if (stage) sc.SG(i - 1) = 0.0;
                    continue;
                }
                ztz = ztz + zi * zi;
                if (stage) sc.SG(i - 1) = zi * scale;
                zi2 = zi1;
                zi1 = zi;
            }
        }
    }
    // downwards from r: this lane covers I = r .. n-1
    {
        double zi = 1.0, zim1 = 0.0;   // Z(I), Z(I-1)
        bool live = act && !nan && r <= n - 1;
        const int loU = wave_min_u((act && !nan) ? r : n);
        if (loU <= n - 1) {
            double um[kUS], ld[kUS], umn[kUS], ldn[kUS];
#pragma unroll
            for (int u = 0; u < kUS; ++u) {
                int i = min(loU + u, n - 1);
                um[u] = sc.UM(i); ld[u] = LD[i - 1];
            }
            for (int i0 = loU; i0 <= n - 1; i0 += kUS) {
                if (!stage && !__any(live)) break;
#pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    int i = min(i0 + kUS + u, n - 1);
                    umn[u] = sc.UM(i); ldn[u] = LD[i - 1];
                }
#pragma unroll
                for (int u = 0; u < kUS; ++u) {
                    const int i = i0 + u;
                    if (act && !nan && i >= r && i <= n - 1) {
                        if (live) {
                            double zn = -(um[u] * zi);
                            if ((fabs(zi) + fabs(zn)) * fabs(ld[u]) < gaptol) {
                                is1 = i;
                                live = false;
                                if (stage) sc.SG(i) = 0.0;
                            } else {
                                ztz = ztz + zn * zn;
                                if (stage) sc.SG(i) = zn * scale;
                                zi = zn;
                            }
                        } else if (stage) {
                            sc.SG(i) = 0.0;
                        }
                    }
                }
#pragma unroll
                for (int u = 0; u < kUS; ++u) { um[u] = umn[u]; ld[u] = ldn[u]; }
            }
        }
        if (nan) {   // slow form after a NaN (odr1v.f:336-350)
            bool cut = false;
            for (int i = r; i <= n - 1; ++i) {
                if (cut) { if (stage) sc.SG(i) = 0.0; continue; }
                double zn;
                if (zi == 0.0) zn = -(LD[i - 2] / LD[i - 1]) * zim1;
                else zn = -(sc.UM(i) * zi);
                if ((fabs(zi) + fabs(zn)) * fabs(LD[i - 1]) < gaptol) {
                    is1 = i;
                    cut = true;
                    if (stage) sc.SG(i) = 0.0;
                    continue;
                }
                ztz = ztz + zn * zn;
                if (stage) sc.SG(i) = zn * scale;
                zim1 = zi;
                zi = zn;
            }
        }
    }
    if (act && !stage) {
        double tmp = 1.0 / ztz;
        o->ztz = ztz;
        o->nrminv = sqrt(tmp);
        o->resid = fabs(o->mingma) * o->nrminv;
        o->rqcorr = o->mingma * tmp;
        o->isupp0 = is0;
        o->isupp1 = is1;
    }
}

// k_singleton: eigenvectors by twisted factorisations with Rayleigh-quotient
// correction (PMR_process_s_task, process_singleton.c:65-324), one lane per
// eigenvector.  Every wave takes a task of up to 64 singletons of one block, so the
// block size, rows and sweep indices are wave-uniform; lanes that share a
// representation (all of them at the root of a block) hit the same cache lines.  Each
// wave owns a
// scratch region of 3 * cap * 64 doubles.  Columns are written once, at the end,
// through an LDS tile so that each store covers contiguous rows of four columns.
__device__ void singleton_group(const STask& t, bool valid, int n, int bbeg,
                                const MatDesc* __restrict__ md, const MatState* __restrict__ ms,
                                const int* __restrict__ Windex, double* W, double* Wsh,
                                double* Werr, const double* __restrict__ Wgap, const WScr& sc,
                                int* supp0, int* supp1) {
    Rrr R;
    R.D = t.D; R.L = t.L; R.LD = t.DL; R.LLD = t.DLL;
    const MatState& S = ms[t.mat];
    const int off = md[t.mat].off;
    const int i = t.slot - off;               // matrix-local eigenvalue index
    const int bl_begin = bbeg, bl_end = bbeg + n - 1;
    const double pivmin = S.pivmin;
    const double tol = 4.0 * log((double)n) * DBL_EPSILON;
    const double RQtol = 2.0 * DBL_EPSILON;
    double lambda = valid ? Wsh[t.slot] : 0.0;
    double werr = valid ? Werr[t.slot] : 0.0;
    double left = lambda - werr, right = lambda + werr;
    const int i_local = valid ? Windex[t.slot] : 1;
    double lgap, rgap;
    if (i == bl_begin) lgap = DBL_EPSILON * fmax(fabs(left), fabs(right));
    else lgap = t.lgap;
    if (i == bl_end) rgap = DBL_EPSILON * fmax(fabs(left), fabs(right));
    else rgap = valid ? Wgap[t.slot] : 0.0;
    const double gap = fmin(lgap, rgap);
    const double gaptol = (i == bl_begin || i == bl_end) ? 0.0 : gap * DBL_EPSILON;
    int usedBS = 0, usedRQ = 0, needBS = 0, wantnc = 1, r = 0;
    bool done = !valid;
    double bstres = 0.0, bstw = 0.0, wsh = lambda;
    R1v o;
    o.r = 1; o.sawnan = 0; o.resid = 0.0; o.rqcorr = 0.0; o.nrminv = 1.0; o.mingma = 0.0;
    o.isupp0 = 1; o.isupp1 = n; o.negcnt = 0; o.ztz = 1.0;
    unsigned long long r1v_el = 0, bis_el = 0, wave_calls = 0;
    for (int kk = 1; kk <= kMaxIterRQI; ++kk) {
        const bool act = !done;
        if (!__any(act)) break;
        ++wave_calls;
        if (act && needBS) {
            usedBS = 1;
            double w = wsh, we = werr;
            bis_el += (unsigned long long)n *
                      dlarrb1(n, t.D, t.DLL, i_local, 0.0, 2.0 * DBL_EPSILON, 0.0, &w, &we, pivmin,
                              t.bl_spdiam, r);
            wsh = w;
            werr = we;
            lambda = wsh;
            r = 0;
        }
        if (act) wantnc = usedBS ? 0 : 1;
        r1v_wave(n, lambda, R, pivmin, wantnc, r, act, sc, &o);
        solve_wave(n, R.LD, gaptol, sc, &o, act, false, 0.0);
        if (act) {
            r1v_el += n;
            r = o.r;
            if (kk == 1 || o.resid < bstres) { bstres = o.resid; bstw = lambda; }
            if (!usedBS && o.resid > tol * gap && fabs(o.rqcorr) > RQtol * fabs(lambda)) {
                int sgndef = (i_local <= o.negcnt) ? -1 : 1;
                if (o.rqcorr * sgndef >= 0.0 && lambda + o.rqcorr <= right &&
                    lambda + o.rqcorr >= left) {
                    usedRQ = 1;
                    if (sgndef == 1) left = lambda;
                    else right = lambda;
                    wsh = 0.5 * (left + right);
                    lambda += o.rqcorr;
                } else needBS = 1;
                if (right - left < RQtol * fabs(lambda)) usedBS = 1;
                else if (kk == kMaxIterRQI - 1) needBS = 1;
            } else done = true;
        }
    }
    const bool step2 = valid && usedRQ && usedBS && bstres <= o.resid;
    if (__any(step2)) {
        ++wave_calls;
        if (step2) lambda = bstw;
        r1v_wave(n, lambda, R, pivmin, wantnc, r, step2, sc, &o);
        solve_wave(n, R.LD, gaptol, sc, &o, step2, false, 0.0);
        if (step2) r1v_el += n;
    }
    solve_wave(n, R.LD, gaptol, sc, &o, valid, true, o.nrminv);
    add_steps(kStepR1v, r1v_el);
    add_steps(kStepR1vBis, bis_el);
    {
        unsigned long long bmax = bis_el;
        for (int o2 = 32; o2 > 0; o2 >>= 1) {
            unsigned long long ot = __shfl_xor(bmax, o2);
            bmax = ot > bmax ? ot : bmax;
        }
        if (threadIdx.x == 0) {
            add_steps(kStepR1vWave, wave_calls * (unsigned long long)n);
            add_steps(kStepBisWave, bmax);
        }
    }
    *supp0 = o.isupp0;
    *supp1 = o.isupp1;
    if (valid) {
        Wsh[t.slot] = lambda;
        Werr[t.slot] = werr;
        W[t.slot] = lambda + t.L[n - 1];
    }
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_singleton(const WaveTask* __restrict__ tasks,
                                                  const LevelCounters* __restrict__ cnt,
                                                  const STask* __restrict__ T,
                                                  const MatDesc* __restrict__ md,
                                                  const MatState* __restrict__ ms,
                                                  const int* __restrict__ Windex,
                                                  const int* __restrict__ Zidx, double* W,
                                                  double* Wsh, double* Werr,
                                                  const double* __restrict__ Wgap, double* scratch,
                                                  int cap) {
    __shared__ double tile[TILE_R][SW + 1];
    const int lane = threadIdx.x;
    const int nT = __builtin_amdgcn_readfirstlane(cnt->nT);
    WScr sc;
    sc.base = scratch + (size_t)blockIdx.x * 3 * (size_t)cap * SW;
    sc.lane = lane;
    for (int ti = blockIdx.x; ti < nT; ti += gridDim.x) {
        const WaveTask wt = tasks[ti];
        const int cntw = __builtin_amdgcn_readfirstlane(wt.cnt);
        const int start = __builtin_amdgcn_readfirstlane(wt.start);
        const bool valid = lane < cntw;
        // lanes past the end replay the last task so that every pointer is valid
        const STask t = T[start + min(lane, cntw - 1)];
        const int n = __builtin_amdgcn_readfirstlane(t.bsize);
        const int bbeg = __builtin_amdgcn_readfirstlane(t.bbeg);
        const MatDesc& dsc = md[t.mat];
        int col = Zidx[t.slot];
        if (col >= dsc.mcols) col = dsc.mcols - 1;   // never write past the caller's Z
        double* Zc = dsc.Z + (size_t)col * dsc.ldz;
        const int nfull = __builtin_amdgcn_readfirstlane(dsc.n);
        int s0 = 1, s1 = 1;
        if (n == 1) {
            sc.SG(0) = 1.0;
        } else {
            singleton_group(t, valid, n, bbeg, md, ms, Windex, W, Wsh, Werr, Wgap, sc, &s0, &s1);
        }
        if (valid) {
            dsc.Zsupp[2 * col] = s0 + bbeg;
            dsc.Zsupp[2 * col + 1] = s1 + bbeg;
        }
        // ---- column write: zero rows outside the block, then the block rows through
        // an LDS tile (16 rows x 4 columns per store)
        for (int c = 0; c < cntw; ++c) {
            double* zc = (double*)__shfl((long long)Zc, c);
            for (int rr = lane; rr < bbeg; rr += SW) zc[rr] = 0.0;
            for (int rr = bbeg + n + lane; rr < nfull; rr += SW) zc[rr] = 0.0;
        }
        const int sub = lane & 15, cq = lane >> 4;
        for (int r0 = 0; r0 < n; r0 += TILE_R) {
#pragma unroll
            for (int rr = 0; rr < TILE_R; ++rr) {
                const int row = min(r0 + rr, n - 1);
                tile[rr][lane] = sc.SG(row);
            }
            __syncthreads();
            for (int cg = 0; cg < SW / 4; ++cg) {
                const int c = cg * 4 + cq;
                double* zc = (double*)__shfl((long long)Zc, c);
                const int row = r0 + sub;
                if (c < cntw && row < n) zc[bbeg + row] = tile[sub][c];
            }
            __syncthreads();
        }
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
                                                  const int* __restrict__ Windex) {
    const int k = blockIdx.x;
    MatState& S = ms[k];
    const MatDesc dsc = md[k];
    if (S.done || S.status != kOk) return;
    const int off = dsc.off, m = S.m;
    const double inv = 1.0 / S.scale;
    const bool wantz = dsc.jobz == kV;
    for (int s = threadIdx.x; s < m; s += blockDim.x) {
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
        dsc.Wout[pos] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (S.scale != 1.0) { S.vl *= inv; S.vu *= inv; }
        if (wantz) {
            // ordering check of sort_eigenpairs (mrrr.c:557-586)
            bool sorted = false;
            while (!sorted) {
                sorted = true;
                for (int i = 0; i < m - 1; ++i) {
                    if (dsc.Wout[i] > dsc.Wout[i + 1]) {
                        sorted = false;
                        double t = dsc.Wout[i]; dsc.Wout[i] = dsc.Wout[i + 1]; dsc.Wout[i + 1] = t;
                        int a = dsc.Zsupp[2 * i]; dsc.Zsupp[2 * i] = dsc.Zsupp[2 * i + 2]; dsc.Zsupp[2 * i + 2] = a;
                        a = dsc.Zsupp[2 * i + 1]; dsc.Zsupp[2 * i + 1] = dsc.Zsupp[2 * i + 3]; dsc.Zsupp[2 * i + 3] = a;
                        double* z0 = dsc.Z + (size_t)i * dsc.ldz;
                        double* z1 = dsc.Z + (size_t)(i + 1) * dsc.ldz;
                        for (int r = 0; r < dsc.n; ++r) { double u = z0[r]; z0[r] = z1[r]; z1[r] = u; }
                    }
                }
            }
        }
        S.done = 1;
    }
}

}  // namespace mr3
