/* A C caller of the drop-in routine, written against the mrrr() prototype that
 * include/mr3hip.h declares (the reference's INCLUDE/mrrr.h:62-65) and linked with
 * -lmr3hip: what a program of the reference's users would do.  Test infrastructure.
 *
 * usage: c_caller IN OUT.  IN holds, little-endian: int n, char jobz, char range,
 * int il, int iu, int tryrac, double vl, double vu, double D[n], double E[n].
 * OUT receives: int info, int m, int tryrac, double vl, double vu, double W[m],
 * double Z[n*m] (column-major, ldz = n) and int Zsupp[2m]. */
#include <stdio.h>
#include <stdlib.h>

#include "mr3hip.h"

static int rd(FILE* f, void* p, size_t sz, size_t cnt) { return fread(p, sz, cnt, f) == cnt; }

int main(int argc, char** argv) {
    if (argc != 3) return 64;
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 65;
    int n, il, iu, tryrac;
    char jobz[2] = {0, 0}, range[2] = {0, 0};
    double vl, vu;
    if (!rd(f, &n, sizeof n, 1) || !rd(f, jobz, 1, 1) || !rd(f, range, 1, 1) || !rd(f, &il, sizeof il, 1) ||
        !rd(f, &iu, sizeof iu, 1) || !rd(f, &tryrac, sizeof tryrac, 1) || !rd(f, &vl, sizeof vl, 1) ||
        !rd(f, &vu, sizeof vu, 1) || n < 1)
        return 66;
    double* D = malloc(sizeof(double) * n);
    double* E = malloc(sizeof(double) * n);
    double* W = malloc(sizeof(double) * n);
    double* Z = malloc(sizeof(double) * (size_t)n * n);
    int* Zsupp = malloc(sizeof(int) * 2 * n);
    if (!D || !E || !W || !Z || !Zsupp || !rd(f, D, sizeof(double), n) || !rd(f, E, sizeof(double), n))
        return 67;
    fclose(f);
    int m = 0, ldz = n;
    int info = mrrr(jobz, range, &n, D, E, &vl, &vu, &il, &iu, &tryrac, &m, W, Z, &ldz, Zsupp);
    FILE* g = fopen(argv[2], "wb");
    if (!g) return 68;
    fwrite(&info, sizeof info, 1, g);
    fwrite(&m, sizeof m, 1, g);
    fwrite(&tryrac, sizeof tryrac, 1, g);
    fwrite(&vl, sizeof vl, 1, g);
    fwrite(&vu, sizeof vu, 1, g);
    if (info == 0 && m > 0) {
        fwrite(W, sizeof(double), m, g);
        if (jobz[0] == 'V' || jobz[0] == 'v') {
            fwrite(Z, sizeof(double), (size_t)n * m, g);
            fwrite(Zsupp, sizeof(int), 2 * (size_t)m, g);
        }
    }
    fclose(g);
    free(D); free(E); free(W); free(Z); free(Zsupp);
    return 0;
}
