// mr3_small.hip -- the reference's route for small matrices on the GPU: one wave per
// matrix runs dstemr (mr3_small_core.h), for every problem of a batch with n < 500,
// which is where the reference's mrrr() calls LAPACK's dstemr (SRC/mrrr.c:149-153).
//
// Layout: the current representation (D, L, D*L, D*L*L) and the work array of dqds,
// dlar1v and dlarrf live in LDS (8 n + 8 doubles, 32 KB at n = 499); the eigenvalue
// bookkeeping (errors, gaps, shifted values, Gershgorin bounds, block and index lists)
// in a per-matrix slice of device memory; W, Z and Zsupp are the caller's arrays.
#pragma once
#include <hip/hip_runtime.h>

#include "mr3_internal.h"

#define SM_HD __host__ __device__
#define SM_SYNC() __syncthreads()
// the route's work into the library's step counters (they count only when profiling
// asks for it; on the host side of a __host__ __device__ function nothing is counted)
#if defined(__HIP_DEVICE_COMPILE__) && !defined(SM_STEPS)
#define SM_STEPS(kind, count) ::mr3::add_steps(::mr3::kStepSmDqds + (kind), (unsigned long long)(count))
#endif
#include "mr3_small_core.h"

namespace mr3 {

constexpr int kSmallBelow = 500;   // DSTEMR_IF_SMALLER (INCLUDE/mrrr.h:158)
constexpr int kSmallNT = 64;       // one wave per matrix

// Threads whose dlar1v / dlarrf work is kept: a node has at most n children.
__host__ __device__ inline int small_work_lanes(int n) { return n < kSmallNT ? n : kSmallNT; }

// Bytes of the per-matrix device slice (mr3sm::carve lays it out).
__host__ __device__ inline long long small_slice_bytes(int n) {
    return mr3sm::slice_bytes(n, small_work_lanes(n));
}

__global__ __launch_bounds__(kSmallNT) void k_small(const MatDesc* __restrict__ md, MatState* ms,
                                                    const int* __restrict__ list,
                                                    const long long* __restrict__ slice_off,
                                                    char* slices) {
    extern __shared__ double lds[];
    __shared__ mr3sm::Shared sh;
    const int k = list[blockIdx.x];
    const MatDesc d = md[k];
    const int n = d.n, tid = threadIdx.x;
    mr3sm::Prob p;
    p.n = n;
    p.wantz = (d.jobz == kV);
    p.cnt = (d.jobz == kC);
    p.range = d.range == kAll ? mr3sm::kAll : (d.range == kVal ? mr3sm::kVal : mr3sm::kInd);
    p.il = d.il; p.iu = d.iu; p.vl = d.vl; p.vu = d.vu;
    p.tryrac = d.tryrac;
    p.ldz = d.ldz;
    p.mcols = d.mcols;
    p.Din = d.Din; p.Ein = d.Ein; p.W = d.Wout; p.Z = d.Z; p.Zsupp = d.Zsupp;
    mr3sm::Work w;
    mr3sm::carve(w, lds, slices + slice_off[blockIdx.x], n, small_work_lanes(n));
    const mr3sm::Result r = mr3sm::dstemr_team(tid, kSmallNT, sh, p, w);
    if (tid == 0) {
        MatState& S = ms[k];
        S.m = r.m;
        S.mtot = r.m;
        S.status = r.info ? kNoRoot : kOk;
        S.tryrac = r.tryrac;
        S.vl = d.vl; S.vu = d.vu; S.il = d.il; S.iu = d.iu;   // dstemr reads them only
        S.done = 1;
    }
}

}  // namespace mr3
