/* C restatement of the oracle's time stepping (forward and reference-mode
 * adjoint), parallel over shot rows with OpenMP.  Built twice: REAL = double
 * (_build/liboracle.so, the FP64 oracle) and REAL = float
 * (_build/liboracle_f32.so, the FP32 mode: IEEE float in the same operation
 * order as the numpy oracle's float32 mode, no contraction, so the two agree
 * bitwise; gradient products are rounded to float and summed in double).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ (cross-checked against the numpy
 * oracle, oracle/fdtd_oracle.py) and by the CPU-baseline leg of bench.py.
 * It takes the per-cell and CPML coefficients from the numpy oracle's
 * set-up and restates the update equations of the reference kernels:
 *   H update ....... src/lib/fields_updates_gpu.cu:59-77
 *   H-CPML ......... src/lib/pml_updates_h.cu:37-53, 153-171, 273-289, 388-404
 *   E update ....... src/lib/fields_updates_gpu.cu:21-32
 *   E-CPML ......... src/lib/pml_updates_e.cu:28-76, 150-168, 267-285, 384-401
 *   source ......... src/lib/sourcereceiver.cu:68-85
 *   receivers ...... src/lib/sourcereceiver.cu:21-35
 *   adjoint source . src/lib/back.cu:24-37
 *   gradient sums .. src/compute.py:1167-1172
 * in the same order as oracle/fdtd_oracle.py (_h_phase, _e_phase, forward,
 * adjoint_reference).
 *
 * Layout: fields (S, NI+1, NJ+1); coefficient planes (NI+1, NJ+1); x-slab
 * memory (S, P, NJ) and y-slab memory (S, NI, P) for E/H x {minus, plus};
 * pml = [ERA, ERB, ERE, ERF, HRA, HRB, HRE, HRF], each of length P.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif

typedef struct {
    int S, NI, NJ, P;
    REAL dx;
    const REAL *uE0, *uE1, *uE4, *uH0, *uH1, *uH4;
    const REAL* pml;
    REAL *Ez, *Hx, *Hy;
    REAL* psi[8];   /* Ex0, Exm, Ey0, Eym, Hx0, Hxm, Hy0, Hym */
} State;

#define F(a, s, i, j) (a)[((size_t)(s) * (st->NI + 1) + (i)) * (st->NJ + 1) + (j)]
#define C(a, i, j) (a)[(size_t)(i) * (st->NJ + 1) + (j)]

static int state_init(State* st, int S, int NI, int NJ, int P, double dx, const REAL* uE0,
                      const REAL* uE1, const REAL* uE4, const REAL* uH0,
                      const REAL* uH1, const REAL* uH4, const REAL* pml) {
    st->S = S; st->NI = NI; st->NJ = NJ; st->P = P; st->dx = (REAL)dx;
    st->uE0 = uE0; st->uE1 = uE1; st->uE4 = uE4;
    st->uH0 = uH0; st->uH1 = uH1; st->uH4 = uH4; st->pml = pml;
    const size_t nf = (size_t)S * (NI + 1) * (NJ + 1);
    st->Ez = calloc(nf, sizeof(REAL));
    st->Hx = calloc(nf, sizeof(REAL));
    st->Hy = calloc(nf, sizeof(REAL));
    int ok = st->Ez && st->Hx && st->Hy;
    for (int k = 0; k < 8; ++k) {
        const size_t n = (k % 4 < 2) ? (size_t)S * P * NJ : (size_t)S * NI * P;
        st->psi[k] = calloc(n > 0 ? n : 1, sizeof(REAL));
        ok = ok && st->psi[k];
    }
    return ok ? 0 : -1;
}

static void state_free(State* st) {
    free(st->Ez); free(st->Hx); free(st->Hy);
    for (int k = 0; k < 8; ++k) free(st->psi[k]);
}

/* H half-step: main update, then the four CPML slabs (x slabs touch Hy,
 * y slabs touch Hx, so their order does not matter here). */
static void h_phase(State* st) {
    const int S = st->S, NI = st->NI, NJ = st->NJ, P = st->P;
    const REAL dx = st->dx, one = (REAL)1, zero = (REAL)0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < S; ++s)
        for (int i = 0; i <= NI; ++i)
            for (int j = 0; j <= NJ; ++j) {
                const REAL ez = F(st->Ez, s, i, j);
                if (i >= 1) {
                    const REAL ezj = j < NJ ? F(st->Ez, s, i, j + 1) : zero;
                    F(st->Hx, s, i, j) = C(st->uH0, i, j) * F(st->Hx, s, i, j) -
                                         C(st->uH1, i, j) * (ezj - ez);
                }
                if (j >= 1) {
                    const REAL ezi = i < NI ? F(st->Ez, s, i + 1, j) : zero;
                    F(st->Hy, s, i, j) = C(st->uH0, i, j) * F(st->Hy, s, i, j) +
                                         C(st->uH1, i, j) * (ezi - ez);
                }
            }
    if (P == 0) return;
    const REAL *RA = st->pml + 4 * P, *RB = st->pml + 5 * P, *RE = st->pml + 6 * P,
                 *RF = st->pml + 7 * P;
    /* x slabs: Hy at rows P-1-q (minus) and NI-P+q (plus) */
#pragma omp parallel for collapse(3) schedule(static)
    for (int side = 0; side < 2; ++side)
        for (int s = 0; s < S; ++s)
            for (int q = 0; q < P; ++q) {
                const int i = side == 0 ? P - 1 - q : NI - P + q;
                REAL* psi = st->psi[4 + side] + ((size_t)s * P + q) * NJ;
                for (int j = 0; j < NJ; ++j) {
                    const REAL d = (F(st->Ez, s, i + 1, j) - F(st->Ez, s, i, j)) / dx;
                    F(st->Hy, s, i, j) = F(st->Hy, s, i, j) +
                                         C(st->uH4, i, j) * ((RA[q] - one) * d + RB[q] * psi[j]);
                    psi[j] = RE[q] * psi[j] - RF[q] * d;
                }
            }
    /* y slabs: Hx at columns P-1-q (minus) and NJ-P+q (plus) */
#pragma omp parallel for collapse(3) schedule(static)
    for (int side = 0; side < 2; ++side)
        for (int s = 0; s < S; ++s)
            for (int i = 0; i < NI; ++i) {
                REAL* psi = st->psi[6 + side] + ((size_t)s * NI + i) * P;
                for (int q = 0; q < P; ++q) {
                    const int j = side == 0 ? P - 1 - q : NJ - P + q;
                    const REAL d = (F(st->Ez, s, i, j + 1) - F(st->Ez, s, i, j)) / dx;
                    F(st->Hx, s, i, j) = F(st->Hx, s, i, j) -
                                         C(st->uH4, i, j) * ((RA[q] - one) * d + RB[q] * psi[q]);
                    psi[q] = RE[q] * psi[q] - RF[q] * d;
                }
            }
}

/* E half-step: main update, then the CPML slabs in the order x0, y0, xmax,
 * ymax (they all touch Ez, so corners see that order). */
static void e_phase(State* st) {
    const int S = st->S, NI = st->NI, NJ = st->NJ, P = st->P;
    const REAL dx = st->dx, one = (REAL)1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < S; ++s)
        for (int i = 1; i <= NI; ++i)
            for (int j = 1; j <= NJ; ++j)
                F(st->Ez, s, i, j) =
                    (C(st->uE0, i, j) * F(st->Ez, s, i, j) +
                     C(st->uE1, i, j) * (F(st->Hy, s, i, j) - F(st->Hy, s, i - 1, j))) -
                    C(st->uE1, i, j) * (F(st->Hx, s, i, j) - F(st->Hx, s, i, j - 1));
    if (P == 0) return;
    const REAL *RA = st->pml, *RB = st->pml + P, *RE = st->pml + 2 * P, *RF = st->pml + 3 * P;
    for (int slab = 0; slab < 4; ++slab) {   /* x0, y0, xmax, ymax */
        const int plus = slab >= 2;
        if (slab % 2 == 0) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int s = 0; s < S; ++s)
                for (int q = 0; q < P; ++q) {
                    const int i = plus ? NI - P + q : P - q;
                    REAL* psi = st->psi[plus] + ((size_t)s * P + q) * NJ;
                    for (int j = 0; j < NJ; ++j) {
                        const REAL d = (F(st->Hy, s, i, j) - F(st->Hy, s, i - 1, j)) / dx;
                        F(st->Ez, s, i, j) = F(st->Ez, s, i, j) +
                                             C(st->uE4, i, j) * ((RA[q] - one) * d + RB[q] * psi[j]);
                        psi[j] = RE[q] * psi[j] - RF[q] * d;
                    }
                }
        } else {
#pragma omp parallel for collapse(2) schedule(static)
            for (int s = 0; s < S; ++s)
                for (int i = 0; i < NI; ++i) {
                    REAL* psi = st->psi[2 + plus] + ((size_t)s * NI + i) * P;
                    for (int q = 0; q < P; ++q) {
                        const int j = plus ? NJ - P + q : P - q;
                        const REAL d = (F(st->Hx, s, i, j) - F(st->Hx, s, i, j - 1)) / dx;
                        F(st->Ez, s, i, j) = F(st->Ez, s, i, j) -
                                             C(st->uE4, i, j) * ((RA[q] - one) * d + RB[q] * psi[q]);
                        psi[q] = RE[q] * psi[q] - RF[q] * d;
                    }
                }
        }
    }
}

/* Forward time loop: sample, H, E, sources (row k of scale), snapshot. */
int oracle_forward(int S, int NI, int NJ, int P, double dx, const REAL* uE0, const REAL* uE1,
                   const REAL* uE4, const REAL* uH0, const REAL* uH1, const REAL* uH4,
                   const REAL* pml, int nsrc, const int64_t* src, const REAL* scale, int nrx,
                   const int64_t* rx, int nt, REAL* traces, REAL* snaps) {
    State stv, *st = &stv;
    if (state_init(st, S, NI, NJ, P, dx, uE0, uE1, uE4, uH0, uH1, uH4, pml)) {
        state_free(st);
        return -1;
    }
    for (int n = 0; n < nt; ++n) {
        for (int s = 0; s < S; ++s)
            for (int r = 0; r < nrx; ++r) {
                const int64_t* c = rx + 2 * ((size_t)s * nrx + r);
                traces[((size_t)s * nt + n) * nrx + r] = F(st->Ez, s, c[0], c[1]);
            }
        h_phase(st);
        e_phase(st);
        for (int k = 0; k < nsrc; ++k)
            for (int s = 0; s < S; ++s) {
                const int64_t* c = src + 2 * ((size_t)s * nsrc + k);
                F(st->Ez, s, c[0], c[1]) = F(st->Ez, s, c[0], c[1]) -
                                           C(st->uE4, c[0], c[1]) * scale[(size_t)k * nt + n];
            }
        if (snaps) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int s = 0; s < S; ++s)
                for (int i = 0; i < NI; ++i)
                    memcpy(snaps + (((size_t)n * S + s) * NI + i) * NJ, &F(st->Ez, s, i, 0),
                           sizeof(REAL) * NJ);
        }
    }
    state_free(st);
    return 0;
}

/* Reference-mode adjoint: for n = nt-1 .. 1 inject -resid at the receivers
 * (in shot-major, receiver order), E, H, then accumulate
 * ge += sum_s (snap[n] - snap[n-1]) * lam, gs += sum_s snap[n] * lam
 * (products in REAL, shots summed in order in double).  ge and gs (NI, NJ,
 * double) are accumulated into. */
int oracle_adjoint(int S, int NI, int NJ, int P, double dx, const REAL* uE0, const REAL* uE1,
                   const REAL* uE4, const REAL* uH0, const REAL* uH1, const REAL* uH4,
                   const REAL* pml, int nrx, const int64_t* rx, int nt, const REAL* resid,
                   const REAL* snaps, double* ge, double* gs) {
    State stv, *st = &stv;
    if (state_init(st, S, NI, NJ, P, dx, uE0, uE1, uE4, uH0, uH1, uH4, pml)) {
        state_free(st);
        return -1;
    }
    for (int n = nt - 1; n >= 1; --n) {
        for (int s = 0; s < S; ++s)
            for (int r = 0; r < nrx; ++r) {
                const int64_t* c = rx + 2 * ((size_t)s * nrx + r);
                F(st->Ez, s, c[0], c[1]) = F(st->Ez, s, c[0], c[1]) -
                                           resid[((size_t)s * nt + n) * nrx + r];
            }
        e_phase(st);
        h_phase(st);
        const REAL* sn = snaps + (size_t)n * S * NI * NJ;
        const REAL* sp = snaps + (size_t)(n - 1) * S * NI * NJ;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                double ae = 0.0, as = 0.0;
                for (int s = 0; s < S; ++s) {
                    const size_t o = ((size_t)s * NI + i) * NJ + j;
                    const REAL lam = F(st->Ez, s, i, j);
                    ae += (double)((sn[o] - sp[o]) * lam);
                    as += (double)(sn[o] * lam);
                }
                ge[(size_t)i * NJ + j] += ae;
                gs[(size_t)i * NJ + j] += as;
            }
    }
    state_free(st);
    return 0;
}

/* Threads the OpenMP runtime will use (for the baseline report). */
int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
