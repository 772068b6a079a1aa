// mr3_internal.h -- records shared by the host driver and the HIP kernels.
//
// Layout in HBM (one launch = one batch of matrices; the drop-in mrrr() is a batch
// of one):
//   * per-row arrays are packed: matrix k owns rows [off_k, off_k + n_k) of D, E,
//     E2, the root representation (D, L = E, DL, DLL), gersch (2 per row), and the
//     eigenvalue slots W, Werr, Wgap, Wsh, Wcoarse, iblock, Windex, Zidx, smat;
//   * block records of matrix k live at [off_k, off_k + nsplit_k) of the block array;
//   * eigenvectors go straight into the caller's Z (column-major, ldz) per matrix.
#pragma once
#include <stdint.h>

namespace mr3 {

// Numerical constants of the reference (INCLUDE/mrrr.h:152-200, mrrr_val.c:95-99,
// process_singleton.c:57, odrrf.f:113-119, odrrd.f:210-212).
constexpr double kMaxGrowth = 64.0;
constexpr double kMinRelgap = 1e-3;
constexpr int kMaxTryRoot = 6;
constexpr double kFudgeRoot = 2.0;
constexpr double kRandFactor = 8.0;
constexpr int kMaxIterRQI = 10;
// steps per chunk of the eigenvector sweeps: checkpoint spacing and LDS buffer rows
constexpr int kChunk = 16;
// spare doubles on both sides of every array a sweep reads (>= the furthest a sweep
// reads past a block end: two prefetch groups plus a chunk)
constexpr int kSweepPad = 64;
constexpr double kMaxGrowth1 = 8.0;
constexpr double kMaxGrowth2 = 8.0;

enum Range : int { kAll = 1, kVal = 2, kInd = 3 };
enum Jobz : int { kN = 0, kV = 1, kC = 2 };

// Status codes (per matrix).
enum Status : int { kOk = 0, kBadInput = 1, kNoRoot = 2, kTooFew = 3 };

struct MatDesc {            // host-filled, one per matrix
    const double* Din;      // caller's diagonal (device)
    const double* Ein;      // caller's off-diagonal (device), n-1 used
    double* Wout;           // caller's eigenvalue output (device), n entries
    double* Z;              // caller's eigenvectors (device), ldz x mcols, may be null
    int* Zsupp;             // caller's support output (device), 2*mcols, may be null
    double vl, vu;
    int n, off;
    int jobz, range;
    int il, iu;
    int tryrac, ldz;
    int mcols, pad;
};

struct MatState {           // device-computed, one per matrix
    double scale, pivmin, spdiam, tnorm;
    double gl, gu;          // fudged Gershgorin bounds of T (mrrr_val.c:239-240)
    double glraw, guraw;    // unfudged bounds
    double wl, wu, wlu, wul;
    double vl, vu;          // interval containing the computed eigenvalues (scaled)
    double rtol1, rtol2, bsrtol, rqtol, split;
    int il, iu, nvals, mtot, m, nsplit, nwl, nwu;
    int tryrac, irange, status, done;
    int maxbl, pad;
};

struct BlockRec {           // one per unreduced block
    int bbeg, bend;         // rows, matrix-local
    int wbeg, wend;         // eigenvalue slots, matrix-local (wend < wbeg: none)
    int ilow, iupp;         // local index range wanted
    int mat, pad;
    double gl, gu;          // mrrr_val block bounds (mrrr_val.c:369-374)
    double glraw, guraw;    // unfudged Gershgorin bounds of the block
    double spdiam;          // spdiam returned by compute_root_rrr
    double sigma;           // root shift
    double bl, bu;          // interval the block's coarse bisection starts from, and
    int nl, nr;             //   the counts at its ends (dlarrd's block loop,
                            //   odrrd.f:455-480): the fudged Gershgorin interval with
                            //   counts 0 and nbl, cut to the block's own range-I bracket
                            //   when only part of its spectrum is wanted
};

// A cluster of the representation tree at depth >= 1 (process_cluster.c cluster_t).
struct ClTask {
    int cb, ce;             // global slot range
    int mat, depth;
    int bbeg, bsize;        // block rows (matrix-local begin) and size
    double bl_spdiam;
    double lgap;
    const double* pD;       // parent representation
    const double* pL;
    const double* pDL;
    const double* pDLL;
    double* cD;             // child representation (arena, 4*bsize doubles)
    double* cL;
    double* cDL;            // also the trial factorisation of the right shift
    double* cDLL;
    double tau;             // shift of the child relative to the parent
    long long arena_off;    // offset of this cluster's slot in the level arena
    int eig_base;           // first entry of this cluster in the level eigen list
    int shift;              // 1: left end, 2: right end (dlarrf SHIFT)
    int gblk;               // global index of the block record
    int wseg;               // first global eigenvalue slot of the block
};

// A wave's share of a per-block list: entries [start, start + cnt), cnt <= 64, all
// of one block (so the block order, rows and loop bounds are wave-uniform).
struct WaveTask {
    int start, cnt;
};

// One eigenvector to compute (process_singleton.c singleton_t with begin == end).
// Rayleigh-quotient state of one singleton between the iteration launches (the loop
// variables of process_singleton.c:150-300).  flags: usedBS | usedRQ << 1 |
// needBS << 2 | wantnc << 3 | state << 4 (active, best-residual step, done).
struct SState {
    double lam, lamv, werr, left, right, wsh, bstres, bstw, nrminv, resid;
    int kk, r, flags, is0, is1, pad;
};

struct STask {
    int slot, mat;
    int bbeg, bsize;
    int gblk, wseg;         // global block index, first eigen slot of the block
    double bl_spdiam, lgap;
    const double* D;
    const double* L;
    const double* DL;
    const double* DLL;
};


// Counters written by the classify kernels and read back by the host once per level.
constexpr int kSizeClasses = 32;   // singleton tasks are ordered by floor(log2(block size))

struct LevelCounters {
    int nS, nC, nEig, maxbl;
    long long arena;        // doubles needed by the next level's cluster slots
    int err, nT;            // nT: wave tasks planned for the singleton kernel
    int next;               // next wave task to hand out
    int nD;                 // singletons in the dense list of the tail kernel (blocks
                            // up to kTailN rows, S and P in LDS)
    int nextD;              // next entry of that list to hand out
    int nDb;                // the same for larger blocks (S and P in device memory)
    int nextDb;
    int cls[kSizeClasses];  // wave tasks per size class
    int clsPos[kSizeClasses];
};

// A started small-matrix solve (the dstemr route, n < 500): small_begin launches it on
// a stream of its own after the work queued on st, small_end waits for it, copies the
// per-problem states back and orders st after it.
struct SmallJob {
    struct Impl;
    Impl* p;
    SmallJob();
    ~SmallJob();
    SmallJob(const SmallJob&) = delete;
    SmallJob& operator=(const SmallJob&) = delete;
};

}  // namespace mr3
