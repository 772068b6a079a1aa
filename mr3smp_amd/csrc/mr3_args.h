// mr3_args.h -- argument checks and problem descriptors shared by the entry points.
#pragma once
#include <cstring>

#include "../../include/mr3hip.h"
#include "mr3_internal.h"

namespace mr3 {

inline int upper(char c) { return (c >= 'a' && c <= 'z') ? c - 32 : c; }

// Argument checks of mrrr.c:118-126.  Returns 0 if valid.
inline int check_args(int jz, int rg, int n, double vl, double vu, int il, int iu) {
    if (!(jz == 'N' || jz == 'V' || jz == 'C')) return 1;
    if (!(rg == 'A' || rg == 'V' || rg == 'I')) return 1;
    if (n <= 0) return 1;
    if (rg == 'V') { if (vu <= vl) return 1; }
    else if (rg == 'I') { if (il < 1 || il > n || iu < il || iu > n) return 1; }
    return 0;
}

// A batch problem's checks: 0 when it runs on the device, -1 when it is answered on
// the host (jobz 'C' with range 'A' or 'I', mrrr.c:130-138: m filled in), 1 when
// invalid.  Sets the scalar outputs to their inputs and m to 0 first.
inline int prepare_problem(mr3_problem& p) {
    p.m = 0;
    p.vl_out = p.vl; p.vu_out = p.vu; p.il_out = p.il; p.iu_out = p.iu;
    p.tryrac_out = p.tryrac;
    p.info = check_args(upper(p.jobz), upper(p.range), p.n, p.vl, p.vu, p.il, p.iu);
    if (p.info) return 1;
    const int jz = upper(p.jobz), rg = upper(p.range);
    if (jz == 'C' && rg != 'V') {
        p.m = (rg == 'A') ? p.n : p.iu - p.il + 1;
        return -1;
    }
    if (jz == 'V' && (p.Z == nullptr || p.Zsupp == nullptr || (p.ldz > 0 && p.ldz < p.n) ||
                      p.mcols < 1)) {
        p.info = 1;
        return 1;
    }
    return 0;
}

inline MatDesc to_desc(const mr3_problem& p) {
    MatDesc d;
    memset(&d, 0, sizeof(d));
    d.Din = p.D; d.Ein = p.E; d.Wout = p.W; d.Z = p.Z; d.Zsupp = p.Zsupp;
    d.vl = p.vl; d.vu = p.vu; d.n = p.n;
    const int jz = upper(p.jobz), rg = upper(p.range);
    d.jobz = jz == 'V' ? kV : (jz == 'N' ? kN : kC);
    d.range = rg == 'A' ? kAll : (rg == 'V' ? kVal : kInd);
    d.il = p.il; d.iu = p.iu; d.tryrac = p.tryrac ? 1 : 0;
    d.ldz = p.ldz > 0 ? p.ldz : p.n;
    d.mcols = p.mcols;
    return d;
}

// Eigenvectors to reserve for a range-V problem: the count of eigenvalues in (vl, vu]
// by the Sturm sequence of T (as dlarrc does with JOBT = 'T', odrrc.f:95-113, which is
// how the reference's dstemr answers the count query) plus a margin of 4 for values
// within rounding of vl or vu, at most n.  The solve reports info 2 if it finds more
// (m > mcols), and mrrr() then solves again with n columns.  Host arrays.
inline int range_v_columns(int n, const double* D, const double* E, double vl, double vu) {
    double lp = D[0] - vl, rp = D[0] - vu;
    int lc = (lp <= 0.0) ? 1 : 0, rc = (rp <= 0.0) ? 1 : 0;
    for (int i = 1; i < n; ++i) {
        const double t = E[i - 1] * E[i - 1];
        lp = (D[i] - vl) - t / lp;
        rp = (D[i] - vu) - t / rp;
        if (lp <= 0.0) ++lc;
        if (rp <= 0.0) ++rc;
    }
    const int c = rc - lc + 4;
    return c < 1 ? 1 : (c > n ? n : c);
}

// A device result into the problem's scalar outputs (rc: the solve's return code).
inline void take_state(mr3_problem& p, const MatState& S, int rc) {
    if (rc != 0) { p.info = 2; return; }
    p.m = S.m;
    p.info = (S.status == kOk) ? 0 : 2;
    p.vl_out = S.vl; p.vu_out = S.vu; p.il_out = S.il; p.iu_out = S.iu;
    p.tryrac_out = S.tryrac;
    if (upper(p.jobz) == 'V' && p.m > p.mcols) p.info = 2;
}

}  // namespace mr3
