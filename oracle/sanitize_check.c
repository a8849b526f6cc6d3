/* Memory-safety check of the C restatement (oracle/fdtd_oracle.c): runs its
 * forward and reference-mode adjoint on small grids -- no CPML, thin and
 * thick CPML, sources and receivers on the domain's corners and padding
 * row/column -- with every array allocated to exactly the size the layout
 * comment of fdtd_oracle.c gives, so that AddressSanitizer reports any read
 * or write outside them and UndefinedBehaviorSanitizer any overflow.  The
 * coefficients are arbitrary finite values: only the indexing is under test.
 *
 * TEST INFRASTRUCTURE ONLY (built by `make -C oracle sanitize`, run by
 * tests/test_c_oracle.py).  Exit status 0 = every case ran. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif

int oracle_forward(int S, int NI, int NJ, int P, double dx, const REAL* uE0, const REAL* uE1,
                   const REAL* uE4, const REAL* uH0, const REAL* uH1, const REAL* uH4,
                   const REAL* pml, int nsrc, const int64_t* src, const REAL* scale, int nrx,
                   const int64_t* rx, int nt, REAL* traces, REAL* snaps);
int oracle_adjoint(int S, int NI, int NJ, int P, double dx, const REAL* uE0, const REAL* uE1,
                   const REAL* uE4, const REAL* uH0, const REAL* uH1, const REAL* uH4,
                   const REAL* pml, int nrx, const int64_t* rx, int nt, const REAL* resid,
                   const REAL* snaps, double* ge, double* gs);

static REAL* filled(size_t n, REAL v) {
    REAL* a = malloc((n ? n : 1) * sizeof(REAL));
    for (size_t k = 0; k < n; ++k) a[k] = v * (REAL)(1 + (k % 7)) / 7;
    return a;
}

static int run_case(int S, int NI, int NJ, int P, int nt, int oob) {
    const size_t plane = (size_t)(NI + 1) * (NJ + 1);
    REAL *c[6];
    const REAL vals[6] = {(REAL)0.9, (REAL)0.3, (REAL)0.2, (REAL)1, (REAL)0.3, (REAL)0.2};
    for (int k = 0; k < 6; ++k) c[k] = filled(plane, vals[k]);
    REAL* pml = filled((size_t)8 * P, (REAL)0.5);
    /* sources: corners of the model and one interior cell; receivers: the
     * same plus the padding row / column (i = NI, j = NJ) */
    const int nsrc = 3, nrx = 6;
    int64_t* src = malloc(sizeof(int64_t) * 2 * S * nsrc);
    int64_t* rx = malloc(sizeof(int64_t) * 2 * S * nrx);
    for (int s = 0; s < S; ++s) {
        const int64_t sp[3][2] = {{0, 0}, {NI - 1, NJ - 1}, {NI / 2, NJ / 3}};
        const int64_t rp[6][2] = {{0, 0}, {NI - 1, NJ - 1}, {NI, NJ}, {NI, 0}, {0, NJ}, {NI / 3, NJ / 2}};
        for (int k = 0; k < nsrc; ++k) { src[2 * (s * nsrc + k)] = sp[k][0]; src[2 * (s * nsrc + k) + 1] = sp[k][1]; }
        for (int k = 0; k < nrx; ++k) { rx[2 * (s * nrx + k)] = rp[k][0]; rx[2 * (s * nrx + k) + 1] = rp[k][1]; }
    }
    if (oob) rx[2 * (S * nrx - 1)] = (int64_t)(S * (NI + 1)) + 1;   /* planted: past the fields */
    REAL* scale = filled((size_t)nsrc * nt, (REAL)1e-3);
    REAL* traces = filled((size_t)S * nt * nrx, 0);
    REAL* snaps = filled((size_t)nt * S * NI * NJ, 0);
    int rc = oracle_forward(S, NI, NJ, P, 0.01, c[0], c[1], c[2], c[3], c[4], c[5], pml, nsrc, src,
                            scale, nrx, rx, nt, traces, snaps);
    double* ge = calloc((size_t)NI * NJ, sizeof(double));
    double* gs = calloc((size_t)NI * NJ, sizeof(double));
    if (!rc)
        rc = oracle_adjoint(S, NI, NJ, P, 0.01, c[0], c[1], c[2], c[3], c[4], c[5], pml, nrx, rx, nt,
                            traces, snaps, ge, gs);
    for (int k = 0; k < 6; ++k) free(c[k]);
    free(pml); free(src); free(rx); free(scale); free(traces); free(snaps); free(ge); free(gs);
    return rc;
}

int main(int argc, char** argv) {
    /* "oob": the planted-mistake run (a receiver past the last shot's
     * fields); AddressSanitizer must stop it */
    const int oob = argc > 1 && argv[1][0] == 'o';
    const int cases[][5] = {{1, 2, 2, 0, 4}, {2, 5, 4, 0, 12}, {2, 12, 10, 2, 20}, {3, 30, 20, 5, 25},
                            {1, 9, 40, 4, 10}, {2, 41, 11, 5, 15}, {1, 3, 3, 1, 6}};
    for (size_t k = 0; k < sizeof(cases) / sizeof(cases[0]); ++k) {
        const int* q = cases[k];
        if (run_case(q[0], q[1], q[2], q[3], q[4], oob)) {
            fprintf(stderr, "case %zu failed\n", k);
            return 1;
        }
    }
    printf("sanitized oracle: %zu cases ran\n", sizeof(cases) / sizeof(cases[0]));
    return 0;
}
