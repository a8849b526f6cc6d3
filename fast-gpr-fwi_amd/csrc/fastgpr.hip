// libfastgpr: 2-D TMz FDTD forward / adjoint with CPML for gfx950.
//
// Update equations follow the reference kernels (file:line in the reference
// repository):
//   H update ....... src/lib/fields_updates_gpu.cu:59-77
//   H-CPML ......... src/lib/pml_updates_h.cu:37-53 (x-), 153-171 (x+),
//                    273-289 (y-), 388-404 (y+)
//   E update ....... src/lib/fields_updates_gpu.cu:21-32
//   E-CPML ......... src/lib/pml_updates_e.cu:28-45/58-76 (x-), 150-168 (x+),
//                    267-285 (y-), 384-401 (y+)
//   source ......... src/lib/sourcereceiver.cu:68-85
//   receivers ...... src/lib/sourcereceiver.cu:21-35
//   adjoint source . src/lib/back.cu:24-37
//   gradient ....... src/compute.py:1167-1172
//   coefficients ... src/lib/uc.cu:11-45, 60-91
//
// Design (MI355X): one thread per (i, j) cell that loops over the shots of a
// batch, so the per-model coefficients are loaded once per cell and reused
// for every shot, and the FP64 gradient accumulation over shots needs no
// atomics.  The CPML correction, source injection, receiver sampling and
// snapshot store are fused into the two half-step kernels.  Implicit
// floating-point contraction is off in this file: the cell updates follow the
// reference's FP32 operation order, with every product that feeds an add or a
// subtract fused explicitly (fmaf, the helpers upd_* / cpml_*), as a CUDA
// build with default flags contracts them, and the CPML derivatives
// (a - b) / dx formed as (a - b) * fl32(1 / dx) (at most one unit in the last
// place apart; an IEEE division costs about ten instructions in the CPML
// cells).  The per-step and the time-blocked kernels share these helpers and
// agree bitwise.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "fastgpr.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                             \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return fail(FGPR_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlockJ = 64;   // one wavefront along a row
constexpr int kBlockI = 4;    // four rows per workgroup (256 threads)

// Slab kinds inside the psi workspace.
enum Slab { EX0 = 0, EXM = 1, HX0 = 2, HXM = 3, EY0 = 4, EYM = 5, HY0 = 6, HYM = 7 };

struct Dev {
    int NI, NJ, P, S, nt, nsrc, nrx, pitch;
    int snap_base, snap_mod;
    int snap_tiled;  // snapshots in the time-blocked layout (tb_snap_off)
    float dx;
    float idx;       // fl32(1 / dx): CPML derivatives are (a - b) * idx
    const float* __restrict__ coef;
    const float* __restrict__ pml;
    const int* __restrict__ src;
    const float* __restrict__ srcscale;
    const int* __restrict__ rx;
    float* Ez;
    float* Hx;
    float* Hy;
    float* psi;
    float* traces;
    float* snaps;
    const float* __restrict__ resid;
    double* grad_eps;
    double* grad_sig;
    const int* ctr;
};

Dev to_dev(const fgpr_problem& p) {
    Dev d;
    d.NI = p.NI; d.NJ = p.NJ; d.P = p.P; d.S = p.S; d.nt = p.nt;
    d.nsrc = p.nsrc; d.nrx = p.nrx; d.pitch = p.pitch; d.dx = p.dx;
    d.idx = 1.0f / p.dx;
    d.snap_base = p.snap_base;
    d.snap_mod = p.snap_mod;
    d.snap_tiled = p.snap_tiled;
    d.coef = p.coef; d.pml = p.pml; d.src = p.src; d.srcscale = p.srcscale; d.rx = p.rx;
    d.Ez = p.Ez; d.Hx = p.Hx; d.Hy = p.Hy; d.psi = p.psi;
    d.traces = p.traces; d.snaps = p.snaps; d.resid = p.resid;
    d.grad_eps = p.grad_eps; d.grad_sig = p.grad_sig; d.ctr = p.ctr;
    return d;
}

__device__ __forceinline__ size_t plane(const Dev& d) { return (size_t)(d.NI + 1) * d.pitch; }

// Offset of snapshot (step n, shot s, cell i, j): slot n - snap_base, or
// n mod snap_mod for a ring of slots (checkpointed segments).
__device__ __forceinline__ size_t snap_off(const Dev& d, int n, int s, int i, int j) {
    const int slot = d.snap_mod ? n % d.snap_mod : n - d.snap_base;
    return (((size_t)slot * d.S + s) * d.NI + i) * d.pitch + j;
}

__device__ __forceinline__ size_t xslab_floats(const Dev& d) { return (size_t)d.S * d.P * d.NJ; }
__device__ __forceinline__ size_t yslab_floats(const Dev& d) { return (size_t)d.S * d.NI * d.P; }

__device__ __forceinline__ float* psi_x(const Dev& d, int kind, int s, int q, int j) {
    return d.psi + (size_t)kind * xslab_floats(d) + ((size_t)s * d.P + q) * d.NJ + j;
}
__device__ __forceinline__ float* psi_y(const Dev& d, int kind, int s, int i, int q) {
    return d.psi + 4 * xslab_floats(d) + (size_t)(kind - 4) * yslab_floats(d) +
           ((size_t)s * d.NI + i) * d.P + q;
}

// Cell arithmetic shared by the per-step and the time-blocked kernels (one
// definition, so both paths round identically).  A product that feeds an add
// or a subtract is fused into one v_fma_f32, as a CUDA build of the
// reference with default flags (-fmad=true) contracts these expressions.
__device__ __forceinline__ float upd_hx(float hx, float c, float d) { return fmaf(-c, d, hx); }  // hx - c d
__device__ __forceinline__ float upd_hy(float hy, float c, float d) { return fmaf(c, d, hy); }   // hy + c d
// (c0 e + c1 dhy) - c1 dhx
__device__ __forceinline__ float upd_e(float c0, float c1, float e, float dhy, float dhx) {
    return fmaf(-c1, dhx, fmaf(c1, dhy, c0 * e));
}
// CPML: correction factor (RA - 1) dd + RB psi and next psi = RE psi - RF dd
__device__ __forceinline__ float cpml_t(float ra01, float rb, float dd, float ps) {
    return fmaf(rb, ps, ra01 * dd);
}
__device__ __forceinline__ float cpml_psi(float re, float rf, float dd, float ps) {
    return fmaf(re, ps, -(rf * dd));
}

// Time index of this launch: absolute `koff` when no counter is bound,
// otherwise (*ctr + dir*koff) so a captured graph can be replayed.
__device__ __forceinline__ int step_index(const Dev& d, int koff, int dir) {
    return d.ctr ? (*d.ctr + dir * koff) : koff;
}

// ---------------------------------------------------------------------------
// Per-cell coefficients (uc.cu:11-45 forward, 60-91 adjoint).
__global__ void k_cell_coeffs(const float* __restrict__ er, const float* __restrict__ se,
                              const float* __restrict__ mr, int NI, int NJ, int pitch,
                              float dt, float dx, int adjoint, float* __restrict__ coef) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= pitch || i > NI) return;
    const float e0 = 8.8541878128e-12f;   // uc.cu:8
    const float m0 = 1.25663706212e-06f;  // uc.cu:9
    float uE0 = 0.f, uE1 = 0.f, uE4 = 0.f, uH1 = 0.f, uH4 = 0.f;
    if (i < NI && j < NJ) {
        const size_t c = (size_t)i * NJ + j;
        const float e = er[c], s = se[c], m = mr[c];
        const float HA = (m0 * m) / dt;
        uH1 = ((1.0f / dx) * 1.0f) / HA;
        uH4 = 1.0f / HA;
        if (!(s > 100.0f)) {
            const float t = (e0 * e) / dt;
            // `(e0*er/dt) + 0.5*se` is evaluated in double and rounded once.
            const float EA = (float)((double)t + 0.5 * (double)s);
            const float EB = (float)((double)t - 0.5 * (double)s);
            if (adjoint) {
                const float a = (2.0f * e0) * e;
                uE0 = a / (a + s * dt);
            } else {
                uE0 = EB / EA;
            }
            uE1 = ((1.0f / dx) * 1.0f) / EA;
            uE4 = 1.0f / EA;
        }
    }
    const size_t pl = (size_t)(NI + 1) * pitch;
    const size_t o = (size_t)i * pitch + j;
    coef[o] = uE0;
    coef[pl + o] = uE1;
    coef[2 * pl + o] = uE4;
    coef[3 * pl + o] = uH1;
    coef[4 * pl + o] = uH4;
}

// ---------------------------------------------------------------------------
// H half-step: H update + H-CPML for all shots of the batch.  With SAMPLE
// the same launch also records the receivers for step n (Ez is read-only in
// this phase, so sampling here equals sampling before the step).
template <bool SAMPLE>
__global__ void __launch_bounds__(kBlockJ* kBlockI)
k_h_phase(Dev d, int koff, int dir) {
    const int n = step_index(d, koff, dir);
    const int j = blockIdx.x * kBlockJ + threadIdx.x;
    const int i = blockIdx.y * kBlockI + threadIdx.y;
    const int NI = d.NI, NJ = d.NJ, P = d.P, pitch = d.pitch;

    if (SAMPLE) {
        const int tid = (blockIdx.y * gridDim.x + blockIdx.x) * (kBlockJ * kBlockI) +
                        threadIdx.y * kBlockJ + threadIdx.x;
        const int nthreads = gridDim.x * gridDim.y * kBlockJ * kBlockI;
        for (int t = tid; t < d.S * d.nrx; t += nthreads) {
            const int s = t / d.nrx, r = t - s * d.nrx;
            const int ri = d.rx[2 * t], rj = d.rx[2 * t + 1];
            float v = 0.f;
            if (ri >= 0) v = d.Ez[((size_t)s * (NI + 1) + ri) * pitch + rj];
            d.traces[((size_t)s * d.nt + n) * d.nrx + r] = v;
        }
    }
    if (i >= NI || j >= NJ) return;

    const size_t pl = plane(d);
    const size_t o = (size_t)i * pitch + j;
    const float uH1 = d.coef[3 * pl + o];
    const float uH4 = d.coef[4 * pl + o];

    // CPML membership of this cell (H slabs sit one cell below the E slabs
    // on the minus side: pml_updates_h.cu:39 vs pml_updates_e.cu:41).
    int xq = -1, xk = 0, yq = -1, yk = 0;
    if (P > 0) {
        if (i < P) { xq = P - 1 - i; xk = HX0; }
        else if (i >= NI - P) { xq = i - (NI - P); xk = HXM; }
        if (j < P) { yq = P - 1 - j; yk = HY0; }
        else if (j >= NJ - P) { yq = j - (NJ - P); yk = HYM; }
    }
    float xRA01 = 0.f, xRB = 0.f, xRE = 0.f, xRF = 0.f;
    float yRA01 = 0.f, yRB = 0.f, yRE = 0.f, yRF = 0.f;
    if (xq >= 0) {
        xRA01 = d.pml[4 * P + xq] - 1.0f; xRB = d.pml[5 * P + xq];
        xRE = d.pml[6 * P + xq]; xRF = d.pml[7 * P + xq];
    }
    if (yq >= 0) {
        yRA01 = d.pml[4 * P + yq] - 1.0f; yRB = d.pml[5 * P + yq];
        yRE = d.pml[6 * P + yq]; yRF = d.pml[7 * P + yq];
    }

    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        const float ez = d.Ez[b];
        const float ezj = d.Ez[b + 1];       // (i, j+1), j+1 <= NJ
        const float ezi = d.Ez[b + pitch];   // (i+1, j), i+1 <= NI
        // Hx: main update for i >= 1 (uH0 == 1), y-CPML on every row.
        if (i >= 1 || yq >= 0) {
            float hx = d.Hx[b];
            if (i >= 1) hx = upd_hx(hx, uH1, ezj - ez);
            if (yq >= 0) {
                float* pp = psi_y(d, yk, s, i, yq);
                const float ps = *pp;
                const float dd = (ezj - ez) * d.idx;
                hx = fmaf(-uH4, cpml_t(yRA01, yRB, dd, ps), hx);
                *pp = cpml_psi(yRE, yRF, dd, ps);
            }
            d.Hx[b] = hx;
        }
        // Hy: main update for j >= 1, x-CPML on every column.
        if (j >= 1 || xq >= 0) {
            float hy = d.Hy[b];
            if (j >= 1) hy = upd_hy(hy, uH1, ezi - ez);
            if (xq >= 0) {
                float* pp = psi_x(d, xk, s, xq, j);
                const float ps = *pp;
                const float dd = (ezi - ez) * d.idx;
                hy = fmaf(uH4, cpml_t(xRA01, xRB, dd, ps), hy);
                *pp = cpml_psi(xRE, xRF, dd, ps);
            }
            d.Hy[b] = hy;
        }
    }
}

// ---------------------------------------------------------------------------
// E half-step: E update + E-CPML (slab order x0, y0, xmax, ymax as
// compute.py:659 launches them), then optionally the source injection and
// snapshot store (forward) or the gradient accumulation (adjoint).
template <bool SRC, bool SNAP, bool GRAD>
__global__ void __launch_bounds__(kBlockJ* kBlockI)
k_e_phase(Dev d, int koff, int dir) {
    const int n = step_index(d, koff, dir);
    const int j = blockIdx.x * kBlockJ + threadIdx.x;
    const int i = blockIdx.y * kBlockI + threadIdx.y;
    const int NI = d.NI, NJ = d.NJ, P = d.P, pitch = d.pitch;
    if (i >= NI || j >= NJ) return;

    const size_t pl = plane(d);
    const size_t o = (size_t)i * pitch + j;
    const float uE0 = d.coef[o];
    const float uE1 = d.coef[pl + o];
    const float uE4 = d.coef[2 * pl + o];

    int xq = -1, xk = 0, yq = -1, yk = 0;
    if (P > 0) {
        if (i >= 1 && i <= P) { xq = P - i; xk = EX0; }
        else if (i >= NI - P) { xq = i - (NI - P); xk = EXM; }
        if (j >= 1 && j <= P) { yq = P - j; yk = EY0; }
        else if (j >= NJ - P) { yq = j - (NJ - P); yk = EYM; }
    }
    float xRA01 = 0.f, xRB = 0.f, xRE = 0.f, xRF = 0.f;
    float yRA01 = 0.f, yRB = 0.f, yRE = 0.f, yRF = 0.f;
    if (xq >= 0) {
        xRA01 = d.pml[xq] - 1.0f; xRB = d.pml[P + xq];
        xRE = d.pml[2 * P + xq]; xRF = d.pml[3 * P + xq];
    }
    if (yq >= 0) {
        yRA01 = d.pml[yq] - 1.0f; yRB = d.pml[P + yq];
        yRE = d.pml[2 * P + yq]; yRF = d.pml[3 * P + yq];
    }
    const bool main_upd = (i >= 1 && j >= 1);
    double acc_e = 0.0, acc_s = 0.0;

    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        float ez = d.Ez[b];
        const float hy = d.Hy[b];
        const float hx = d.Hx[b];
        const float hyim = (i >= 1) ? d.Hy[b - pitch] : 0.f;
        const float hxjm = (j >= 1) ? d.Hx[b - 1] : 0.f;
        if (main_upd) ez = upd_e(uE0, uE1, ez, hy - hyim, hx - hxjm);
        // x-CPML corrects Ez through dHy/dx, y-CPML through dHx/dy.
        float xcorr_d = 0.f, ycorr_d = 0.f;
        if (xq >= 0) xcorr_d = (hy - hyim) * d.idx;
        if (yq >= 0) ycorr_d = (hx - hxjm) * d.idx;
        if (xk == EX0 && xq >= 0) {
            float* pp = psi_x(d, EX0, s, xq, j);
            const float ps = *pp;
            ez = fmaf(uE4, cpml_t(xRA01, xRB, xcorr_d, ps), ez);
            *pp = cpml_psi(xRE, xRF, xcorr_d, ps);
        }
        if (yk == EY0 && yq >= 0) {
            float* pp = psi_y(d, EY0, s, i, yq);
            const float ps = *pp;
            ez = fmaf(-uE4, cpml_t(yRA01, yRB, ycorr_d, ps), ez);
            *pp = cpml_psi(yRE, yRF, ycorr_d, ps);
        }
        if (xk == EXM && xq >= 0) {
            float* pp = psi_x(d, EXM, s, xq, j);
            const float ps = *pp;
            ez = fmaf(uE4, cpml_t(xRA01, xRB, xcorr_d, ps), ez);
            *pp = cpml_psi(xRE, xRF, xcorr_d, ps);
        }
        if (yk == EYM && yq >= 0) {
            float* pp = psi_y(d, EYM, s, i, yq);
            const float ps = *pp;
            ez = fmaf(-uE4, cpml_t(yRA01, yRB, ycorr_d, ps), ez);
            *pp = cpml_psi(yRE, yRF, ycorr_d, ps);
        }
        if (SRC) {
            for (int k = 0; k < d.nsrc; ++k) {
                const int* sc = d.src + 2 * ((size_t)s * d.nsrc + k);
                if (sc[0] == i && sc[1] == j) ez = fmaf(-uE4, d.srcscale[(size_t)k * d.nt + n], ez);
            }
        }
        d.Ez[b] = ez;
        if (SNAP) d.snaps[snap_off(d, n, s, i, j)] = ez;
        if (GRAD) {
            const size_t sn = snap_off(d, n, s, i, j);
            const size_t sp = snap_off(d, n - 1, s, i, j);
            const float fn = d.snaps[sn];
            const float fp = d.snaps[sp];
            acc_e += (double)((fn - fp) * ez);
            acc_s += (double)(fn * ez);
        }
    }
    if (GRAD) {
        const size_t g = (size_t)i * NJ + j;
        d.grad_eps[g] += acc_e;
        d.grad_sig[g] += acc_s;
    }
}

// Adjoint source: lambda_Ez[s, rx] -= resid[s, n, r] (back.cu:24-37).  A
// receiver on i = NI or j = NJ is skipped: the E update zeroes that cell
// before anything reads it.
// Receivers on one cell are applied by the thread of the first of them, one
// after another in list order (Ez - g_a - g_b ...), as the time-blocked
// kernels do: the result does not depend on thread timing (the reference's
// back.cu:33-34 races on such cells).
__device__ __forceinline__ bool first_on_cell(const Dev& d, int s, int r, int ri, int rj) {
    const int* rs = d.rx + 2 * (size_t)s * d.nrx;
    for (int q = 0; q < r; ++q)
        if (rs[2 * q] == ri && rs[2 * q + 1] == rj) return false;
    return true;
}

__global__ void k_adj_inject(Dev d, int koff, int dir) {
    const int n = step_index(d, koff, dir);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.S * d.nrx) return;
    const int s = t / d.nrx, r = t - s * d.nrx;
    const int ri = d.rx[2 * t], rj = d.rx[2 * t + 1];
    if (ri < 0 || ri >= d.NI || rj >= d.NJ || !first_on_cell(d, s, r, ri, rj)) return;
    const int* rs = d.rx + 2 * (size_t)s * d.nrx;
    const float* g = d.resid + ((size_t)s * d.nt + n) * d.nrx;
    float* e = &d.Ez[((size_t)s * (d.NI + 1) + ri) * d.pitch + rj];
    float v = *e;
    for (int q = r; q < d.nrx; ++q)
        if (rs[2 * q] == ri && rs[2 * q + 1] == rj) v = v - g[q];
    *e = v;
}

// ---------------------------------------------------------------------------
// Exact discrete adjoint: the transpose of the forward time stepping, for
// gradients that finite differences of the forward map can check.  Reverse
// step n takes the adjoints of (E^{n+1}, H^{n+1/2}, psi) to those of
// (E^n, H^{n-1/2}, psi) in four launches:
//   A: per E cell -- gradient terms, the E-CPML transposes (Lx, Ly, in the
//      slabs) and the psi_E adjoints;
//   B: per H cell -- gather U = uE1 * lambda_E (formed from lambda_E) and
//      Lx / Ly into lambda_H, then the H-CPML transposes (Mx, My, in the
//      slabs) and the psi_H adjoints;
//   C: per E cell -- lambda_E = (uE0 or 1) lambda + gathers of
//      V = uH1 * lambda_H (formed from lambda_H) and M;
//   D: lambda_E at the receivers += dL/dtraces[n].
// The scratch holds the four slab planes Lx, Ly, Mx, My; every entry read is
// written earlier in the same step, so it needs no initialisation.
// dE^{n+1}/d er = -(e0/dt) uE4 (E^{n+1} - E^n) and, on cells the main update
// covers, dE^{n+1}/d se = -uE4 (E^{n+1} + E^n)/2 (every other term of the E
// update is uE4 or uE1 = uE4/dx times something independent of the cell's
// own er / se).  grad_eps accumulates sum uE4*lambda*(E^{n+1}-E^n); the
// constant -(e0/dt) is applied by the caller.
struct XScratch {
    float* Lx; float* Ly; float* Mx; float* My;
};

// Whether E row / column q (0 <= q < n) lies in a CPML slab: [1, P] or
// [n - P, n) (the E slab indices of k_xadj_a, per axis).
__device__ __forceinline__ bool xslab_e(int q, int n, int P) {
    return P > 0 && ((q >= 1 && q <= P) || q >= n - P);
}

// The same for H rows / columns: [0, P) or [n - P, n).
__device__ __forceinline__ bool xslab_h(int q, int n, int P) {
    return P > 0 && (q < P || q >= n - P);
}

__device__ __forceinline__ XScratch xscratch(const Dev& d, float* base) {
    const size_t f = (size_t)d.S * plane(d);
    XScratch x;
    x.Lx = base; x.Ly = base + f; x.Mx = base + 2 * f; x.My = base + 3 * f;
    return x;
}

__global__ void __launch_bounds__(kBlockJ* kBlockI)
k_xadj_a(Dev d, int koff, int dir, float* scratch) {
    const int n = step_index(d, koff, dir);
    const int j = blockIdx.x * kBlockJ + threadIdx.x;
    const int i = blockIdx.y * kBlockI + threadIdx.y;
    const int NI = d.NI, NJ = d.NJ, P = d.P, pitch = d.pitch;
    if (i >= NI || j >= NJ) return;
    const XScratch x = xscratch(d, scratch);
    const size_t pl = plane(d);
    const size_t o = (size_t)i * pitch + j;
    const float uE4 = d.coef[2 * pl + o];
    const bool mainE = (i >= 1 && j >= 1);
    int xq = -1, xk = 0, yq = -1, yk = 0;
    if (P > 0) {
        if (i >= 1 && i <= P) { xq = P - i; xk = EX0; }
        else if (i >= NI - P) { xq = i - (NI - P); xk = EXM; }
        if (j >= 1 && j <= P) { yq = P - j; yk = EY0; }
        else if (j >= NJ - P) { yq = j - (NJ - P); yk = EYM; }
    }
    double acc_e = 0.0, acc_s = 0.0;
    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        const float e = d.Ez[b];
        const double f1 = d.snaps[snap_off(d, n, s, i, j)];
        const double f0 = n >= 1 ? (double)d.snaps[snap_off(d, n - 1, s, i, j)] : 0.0;
        const double ue = (double)uE4 * (double)e;
        acc_e += ue * (f1 - f0);
        acc_s += mainE ? -0.5 * ue * (f1 + f0) : -0.5 * ue * (f1 - f0);
        // U = uE1 * lambda_E is formed where it is used (k_xadj_b); Lx / Ly
        // exist in the CPML slabs only and are read only there
        if (xq >= 0) {
            float* pm = psi_x(d, xk, s, xq, j);
            const float m = *pm;
            x.Lx[b] = uE4 * (d.pml[xq] - 1.0f) * e - d.pml[3 * P + xq] * m;
            *pm = uE4 * d.pml[P + xq] * e + d.pml[2 * P + xq] * m;
        }
        if (yq >= 0) {
            float* pm = psi_y(d, yk, s, i, yq);
            const float m = *pm;
            x.Ly[b] = -uE4 * (d.pml[yq] - 1.0f) * e - d.pml[3 * P + yq] * m;
            *pm = -uE4 * d.pml[P + yq] * e + d.pml[2 * P + yq] * m;
        }
    }
    const size_t g = (size_t)i * NJ + j;
    d.grad_eps[g] += acc_e;
    d.grad_sig[g] += acc_s;
}

__global__ void __launch_bounds__(kBlockJ* kBlockI)
k_xadj_b(Dev d, float* scratch) {
    const int j = blockIdx.x * kBlockJ + threadIdx.x;
    const int i = blockIdx.y * kBlockI + threadIdx.y;
    const int NI = d.NI, NJ = d.NJ, P = d.P, pitch = d.pitch;
    if (i >= NI || j >= NJ) return;
    const XScratch x = xscratch(d, scratch);
    const size_t pl = plane(d);
    const size_t o = (size_t)i * pitch + j;
    const float uH4 = d.coef[4 * pl + o];
    const float idx = d.idx;
    int xq = -1, xk = 0, yq = -1, yk = 0;
    if (P > 0) {
        if (i < P) { xq = P - 1 - i; xk = HX0; }
        else if (i >= NI - P) { xq = i - (NI - P); xk = HXM; }
        if (j < P) { yq = P - 1 - j; yk = HY0; }
        else if (j >= NJ - P) { yq = j - (NJ - P); yk = HYM; }
    }
    // U = uE1 * lambda_E of the E cells (i, j), (i + 1, j), (i, j + 1): zero
    // off the main E update (row 0, column 0) and outside the grid; Lx / Ly
    // of the same cells: zero outside the E slabs (k_xadj_a writes them
    // there only)
    const float e1c = d.coef[pl + o];
    const float e1d = d.coef[pl + o + pitch];
    const float e1r = d.coef[pl + o + 1];
    const bool uc = i >= 1 && j >= 1, ud = i + 1 < NI && j >= 1, ur = j + 1 < NJ && i >= 1;
    const bool lxc = xslab_e(i, NI, P), lxd = i + 1 < NI && xslab_e(i + 1, NI, P);
    const bool lyc = xslab_e(j, NJ, P), lyr = j + 1 < NJ && xslab_e(j + 1, NJ, P);
    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        const float u0 = uc ? e1c * d.Ez[b] : 0.f;
        const float u_d = ud ? e1d * d.Ez[b + pitch] : 0.f;
        const float u_r = ur ? e1r * d.Ez[b + 1] : 0.f;
        const float lx0 = lxc ? x.Lx[b] : 0.f, lx_d = lxd ? x.Lx[b + pitch] : 0.f;
        const float ly0 = lyc ? x.Ly[b] : 0.f, ly_r = lyr ? x.Ly[b + 1] : 0.f;
        float hy = d.Hy[b] + (u0 - u_d) + (lx0 - lx_d) * idx;
        float hx = d.Hx[b] + (u_r - u0) + (ly0 - ly_r) * idx;
        // Hx on row 0 and Hy on column 0 belong to the closed, identically
        // zero boundary subsystems of the forward map (only CPML terms update
        // them); their adjoints never reach the interior, have a zero
        // gradient weight, and the subsystem is not stable backwards, so
        // they are held at zero.
        if (i == 0) hx = 0.f;
        if (j == 0) hy = 0.f;
        d.Hy[b] = hy;
        d.Hx[b] = hx;
        // Mx / My in the H slabs only (read only there); V = uH1 * lambda_H is
        // formed where it is used (k_xadj_c)
        if (xq >= 0) {
            float* pm = psi_x(d, xk, s, xq, j);
            const float m = *pm;
            x.Mx[b] = uH4 * (d.pml[4 * P + xq] - 1.0f) * hy - d.pml[7 * P + xq] * m;
            *pm = uH4 * d.pml[5 * P + xq] * hy + d.pml[6 * P + xq] * m;
        }
        if (yq >= 0) {
            float* pm = psi_y(d, yk, s, i, yq);
            const float m = *pm;
            x.My[b] = -uH4 * (d.pml[4 * P + yq] - 1.0f) * hx - d.pml[7 * P + yq] * m;
            *pm = -uH4 * d.pml[5 * P + yq] * hx + d.pml[6 * P + yq] * m;
        }
    }
}

__global__ void __launch_bounds__(kBlockJ* kBlockI)
k_xadj_c(Dev d, float* scratch) {
    const int j = blockIdx.x * kBlockJ + threadIdx.x;
    const int i = blockIdx.y * kBlockI + threadIdx.y;
    const int NI = d.NI, NJ = d.NJ, pitch = d.pitch;
    if (i >= NI || j >= NJ) return;
    const XScratch x = xscratch(d, scratch);
    const size_t pl = plane(d);
    const size_t o = (size_t)i * pitch + j;
    const float c0 = (i >= 1 && j >= 1) ? d.coef[o] : 1.0f;
    const float idx = d.idx;
    const int P = d.P;
    // Vx = uH1 * lambda_Hx (zero on row 0), Vy = uH1 * lambda_Hy (zero on
    // column 0) of the H cells (i, j), (i, j - 1), (i - 1, j); Mx / My: zero
    // outside the H slabs
    const float h1c = d.coef[3 * pl + o];
    const float h1l = j >= 1 ? d.coef[3 * pl + o - 1] : 0.f;
    const float h1u = i >= 1 ? d.coef[3 * pl + o - pitch] : 0.f;
    const bool mxc = xslab_h(i, NI, P), mxu = i >= 1 && xslab_h(i - 1, NI, P);
    const bool myc = xslab_h(j, NJ, P), myl = j >= 1 && xslab_h(j - 1, NJ, P);
    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        const float vx = i >= 1 ? h1c * d.Hx[b] : 0.f;
        const float vy = j >= 1 ? h1c * d.Hy[b] : 0.f;
        const float mx = mxc ? x.Mx[b] : 0.f, my = myc ? x.My[b] : 0.f;
        float v = c0 * d.Ez[b] + vx - vy - mx * idx - my * idx;
        if (j >= 1) {
            const float vx_l = i >= 1 ? h1l * d.Hx[b - 1] : 0.f;
            const float my_l = myl ? x.My[b - 1] : 0.f;
            v = v - vx_l + my_l * idx;
        }
        if (i >= 1) {
            const float vy_u = j >= 1 ? h1u * d.Hy[b - pitch] : 0.f;
            const float mx_u = mxu ? x.Mx[b - pitch] : 0.f;
            v = v + vy_u + mx_u * idx;
        }
        d.Ez[b] = (i >= 1 && j >= 1) ? v : 0.f;   // boundary subsystem, see k_xadj_b
    }
}

// lambda_E at the receivers += dL/dtraces[n] (receivers on i = NI or j = NJ
// record a field that is identically zero: nothing to inject).
__global__ void k_xadj_record(Dev d, int koff, int dir) {
    const int n = step_index(d, koff, dir);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.S * d.nrx) return;
    const int s = t / d.nrx, r = t - s * d.nrx;
    const int ri = d.rx[2 * t], rj = d.rx[2 * t + 1];
    if (ri < 0 || ri >= d.NI || rj >= d.NJ || !first_on_cell(d, s, r, ri, rj)) return;
    const int* rs = d.rx + 2 * (size_t)s * d.nrx;
    const float* g = d.resid + ((size_t)s * d.nt + n) * d.nrx;
    float* e = &d.Ez[((size_t)s * (d.NI + 1) + ri) * d.pitch + rj];
    float v = *e;
    for (int q = r; q < d.nrx; ++q)      // receivers on one cell in list order
        if (rs[2 * q] == ri && rs[2 * q + 1] == rj) v = v + g[q];
    *e = v;
}

__global__ void k_advance_ctr(int* ctr, int delta) {
    if (threadIdx.x == 0) ctr[0] = ctr[0] + delta;
}

__global__ void k_set_ctr(int* ctr, int v) {
    if (threadIdx.x == 0) ctr[0] = v;
}

// Deterministic single-workgroup sum of squares, then scaled FP32 output.
__global__ void __launch_bounds__(1024) k_normalise(const double* __restrict__ g, int64_t n,
                                                    double scale, int normalise,
                                                    float* __restrict__ out) {
    __shared__ double red[1024];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += 1024) acc += g[k] * g[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double inv = normalise ? scale / sqrt(red[0]) : scale;
    for (int64_t k = threadIdx.x; k < n; k += 1024) out[k] = (float)(g[k] * inv);
}

// ---------------------------------------------------------------------------
// Time-blocked kernels.
//
// A workgroup owns a TI x TJ tile of one shot and advances it up to K steps
// per launch.  It loads the tile plus a halo -- a region of RI rows x 64
// columns -- of Ez, Hx, Hy and the CPML memory variables from the "in"
// buffers, runs the steps on chip, and writes the tile centre to the "out"
// buffers (ping-pong, so neighbouring workgroups never read cells another one
// is writing).  Each step the region's interior edges go stale by one more
// cell; the halo of K cells on the low side and K + 1 on the high side (H
// reads Ez at +1) absorbs that.  Domain edges are exact: cells outside the
// domain read as zero and are never updated.
//
// Lane = column, wave = a block of RI / W consecutive rows.  A thread keeps
// its cells' fields and the three per-cell coefficients (cE0, cE1, cH1) in
// registers for the whole launch; the column neighbours come from the
// adjacent lane through DPP wave shifts, the row neighbours from the same
// thread except at the wave's first / last row, which are exchanged through
// two small LDS rows per wave.  Only the CPML memory variables (one LDS plane
// per slab kind over the region), the CPML coefficient table and the
// receiver lists live in LDS; cE4 / cH4 are read from the (L2-resident)
// coefficient planes where a CPML cell or a source needs them.
constexpr int kTBCols = 64;
constexpr int kPmax = 128;
constexpr int kWaveList = 128;   // receiver entries one wave can hold
constexpr unsigned kNoEntry = 0xFFFFFFFFu;   // empty slot of a receiver block

// Snapshot block of one (slot, shot, tile) in the time-blocked layout of
// variant (RI, NT, K): waves kWaveLo .. kWaveHi are those that own a
// tile-centre row (rows K .. K + TI - 1 of the region), each storing its CPT
// rows for the TJ centre lanes.
template <int RI, int NT, int K>
struct TBSnapGeom {
    static constexpr int W = NT / 64, CPT = RI / W;
    static constexpr int TI = RI - 2 * K - 1, TJ = kTBCols - 2 * K - 1;
    static constexpr int kWaveLo = K / CPT, kWaveHi = (K + TI - 1) / CPT;
    static constexpr int kBlock = (kWaveHi - kWaveLo + 1) * TJ * CPT;
};

// The same block size for a run-time variant.
__host__ __device__ inline int64_t tb_snap_block(int ri, int nt, int k) {
    const int cpt = ri / (nt / 64), TI = ri - 2 * k - 1, TJ = kTBCols - 2 * k - 1;
    return (int64_t)((k + TI - 1) / cpt - k / cpt + 1) * TJ * cpt;
}

// Exact unsigned division by a launch-constant d through a multiply and a
// shift (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant with the add): for
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (umulhi(n, m) + n) >> l for every 0 <= n < 2^31.  Replaces the
// ~20-instruction integer-division sequences of the work-item mapping.
struct FastDiv {
    unsigned m;
    int l;
};

inline FastDiv make_fastdiv(int d) {
    if (d < 1) d = 1;   // divisors that can be 0 are never divided by (see tb_item)
    int l = 0;
    while ((1u << l) < (unsigned)d) ++l;
    FastDiv f;
    f.m = (unsigned)(((uint64_t)1 << 32) * (((uint64_t)1 << l) - (uint64_t)d) / (uint64_t)d + 1);
    f.l = l;
    return f;
}

// Late reads of rarely used launch arguments (receiver and source tables,
// output pointers).  `reload` returns the address the arguments live at
// (the kernel-argument segment, or a multi-problem table entry) with its
// provenance hidden from the optimiser by an empty asm (no instruction, no
// memory access), and `ld_arg` reads one field through the constant address
// space (a scalar load) where it is used, so these fields are not held in
// scalar registers (and spilled) across the whole kernel.
template <class T>
__device__ __forceinline__ const T* reload(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <class T>
__device__ __forceinline__ T ld_arg(const T* field) {
    return *(const __attribute__((address_space(4))) T*)field;
}

__host__ __device__ __forceinline__ int fdiv(int n, FastDiv f) {
#if defined(__HIP_DEVICE_COMPILE__)
    const unsigned t = __umulhi((unsigned)n, f.m);
#else
    const unsigned t = (unsigned)(((uint64_t)(unsigned)n * f.m) >> 32);   // host copy (tests)
#endif
    return (int)(((uint64_t)t + (unsigned)n) >> f.l);
}

struct TBArgs {
    Dev d;
    const float* Ez_in;
    const float* Hx_in;
    const float* Hy_in;
    const float* psi_in;
    float* Ez_out;
    float* Hx_out;
    float* Hy_out;
    float* psi_out;
    int ti_count, tj_count;
    int int_ti0, int_ni, int_tj0, int_nj;   // rectangle of interior tiles (no CPML cell, no edge)
    FastDiv fd_S, fd_nj, fd_per, fd_tj;     // S, int_nj, tj_count - int_nj, tj_count
    const unsigned* rx_list;   // per (shot, tile) list_cap entries: receivers sampled (forward) /
                               // injected (adjoint), li << 26 | lj << 20 | r, then kNoEntry
    int list_cap;
    double* gshot;       // adjoint: S x 2 x NI x NJ per-shot partial gradients (shot-major)
};

// Work order of a launch.  Workgroups are dispatched round-robin over the 8
// XCDs (block b to XCD b % 8, in order b / 8 within it).  Each XCD gets an
// equal share of the interior items first and then an equal share of the
// general (CPML / edge) items, which cost more (their waves run the CPML and
// edge copies of the step loop): the cheap ones finish early and free their
// CUs for the expensive ones, instead of an expensive tail.  An item is (tile, shot), shots innermost, so the shots of
// a tile (which share coefficients) meet in one XCD's L2.  Sets the shot
// and the tile of block `wg` (a bijection over [0, nwg)) and returns whether
// the item is one of the interior ones.  Compiled for the host too
// (fgpr_tb_map), so its arithmetic can be checked without a GPU.
__host__ __device__ __forceinline__ bool tb_item(const TBArgs& a, int S, int wg, int nwg, int& s,
                                                 int& tile) {
    const int tjc = a.tj_count;
    const int ni = a.int_ni, nj = a.int_nj;
    const int Ni = ni * nj * S;                      // interior items
    const int x = wg & 7, t = wg >> 3;
    const int q = nwg >> 3, r = nwg & 7;
    const int qi = Ni >> 3, ri = Ni & 7;
    const int base = x * q + (x < r ? x : r);        // items of XCDs < x
    const int ia = x * qi + (x < ri ? x : ri);       // interior items of XCDs < x
    const int icnt = qi + (x < ri ? 1 : 0);
    if (t < icnt) {
        const int k = ia + t, kt = fdiv(k, a.fd_S);
        s = k - kt * S;
        const int q2 = fdiv(kt, a.fd_nj);
        const int ti = a.int_ti0 + q2;
        tile = ti * tjc + a.int_tj0 + (kt - q2 * nj);
        return true;
    }
    const int g = (base - ia) + (t - icnt), gt = fdiv(g, a.fd_S);
    s = g - gt * S;
    // general tiles in row-major order with the interior rectangle skipped
    const int before = a.int_ti0 * tjc;              // rows above the rectangle
    const int per = tjc - nj;                        // general tiles of a rectangle row
    if (ni == 0 || gt < before) { tile = gt; return false; }
    const int u = gt - before;
    if (u < ni * per) {
        const int q3 = fdiv(u, a.fd_per);
        const int ti = a.int_ti0 + q3, c = u - q3 * per;
        tile = ti * tjc + (c < a.int_tj0 ? c : c + nj);
        return false;
    }
    tile = (a.int_ti0 + ni) * tjc + (u - ni * per);
    return false;
}

// Device buffers are global memory.  Pointers read from a launch table (the
// multi-problem kernel) carry no address space, and the compiler would reach
// them with flat instructions, which also count against the LDS counter: every
// LDS wait -- each barrier -- would then wait for the outstanding global loads
// and stores as well.  The time-blocked kernels access their buffers through
// this cast instead.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}
typedef float fvec4 __attribute__((ext_vector_type(4)));   // 16-byte vector accesses

// Load through the scalar cache (a wave-uniform address of read-only data).
__device__ __forceinline__ int sload(const int* p) {
    return *(const __attribute__((address_space(4))) int*)p;
}

__device__ __forceinline__ float lane_next(float v) {   // value of lane + 1 (0 on lane 63)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_prev(float v) {   // value of lane - 1 (0 on lane 0)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v),
                                                                  0x138, 0xf, 0xf, false));
}

template <int RI, int NT, int K, bool ADJ>
__device__ __forceinline__ void tb_body(const TBArgs& a, const TBArgs* ap, const int n0,
                                        const int nsteps, const int wg, const int nwg) {
    constexpr int W = NT / 64;
    constexpr int CPT = RI / W;          // rows owned by one thread (a contiguous block)
    constexpr int TI = RI - 2 * K - 1;
    constexpr int TJ = kTBCols - 2 * K - 1;
    static_assert(RI % W == 0 && TI > 0 && TJ > 0 && CPT <= 16, "bad time-blocking geometry");
    const Dev& d = a.d;
    const int NI = d.NI, NJ = d.NJ, P = d.P, pitch = d.pitch;

    const int ntiles = a.ti_count * a.tj_count;
    int s, tile;
    tb_item(a, d.S, wg, nwg, s, tile);
    const int ti = fdiv(tile, a.fd_tj), tj = tile - ti * a.tj_count;
    const int gi0 = ti * TI - K, gj0 = tj * TJ - K;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int r0 = w * CPT;              // first region row of this wave
    const int gj = gj0 + lane;

    // Rows meet between waves only: each wave publishes its first Ez row and
    // last Hy row; columns meet inside a wave through lane shifts.
    __shared__ float sEzTop[W * 64];
    __shared__ float sHyBot[W * 64];
    __shared__ float sPsi[4][RI * 64];   // CPML memory of the region: Hx-slab, Ex-slab, Hy-slab, Ey-slab
    __shared__ int sRx[W][kWaveList];    // receivers of each wave: (m << 26) | (lane << 20) | r

    const size_t pl = plane(d);
    const size_t shot_off = (size_t)s * pl;
    const bool colx = (gj >= 0 && gj <= NJ);
    const bool colact = (gj >= 0 && gj < NJ);
    const bool centre_col = lane >= K && lane < K + TJ;

    int yqh = -1, yqe = -1;
    bool yh_lo = true, ye_lo = true;
    if (P > 0 && colact) {
        if (gj < P) yqh = P - 1 - gj;
        else if (gj >= NJ - P) { yqh = gj - (NJ - P); yh_lo = false; }
        if (gj >= 1 && gj <= P) yqe = P - gj;
        else if (gj >= NJ - P) { yqe = gj - (NJ - P); ye_lo = false; }
    }

    float ez[CPT], hx[CPT], hy[CPT];
    float cE0[CPT], cE1[CPT], cE4[CPT], cH1[CPT], cH4[CPT];
    unsigned act = 0u;   // bit m: cell m is updated (0 <= i < NI, 0 <= j < NJ)

    // x-slab index of row gi for H (one cell lower on the minus side) and E.
    auto xq_h = [&](int gi) -> int {
        return gi < P ? P - 1 - gi : (gi >= NI - P ? gi - (NI - P) : -1);
    };
    auto xq_e = [&](int gi) -> int {
        return (gi >= 1 && gi <= P) ? P - gi : (gi >= NI - P ? gi - (NI - P) : -1);
    };
    const auto* const pin = gmem(a.psi_in);
    const size_t xs = xslab_floats(d), ys = yslab_floats(d);

    // Set-up: every load of the region depends only on the launch arguments
    // and the work item -- the fields, coefficients and CPML memory, this
    // (shot, tile)'s receiver block (self-contained entries: no load waits on
    // another) and the CPML coefficient table -- so all of them are in
    // flight together, one memory round trip.
    // A wave whose rows hold no CPML row, no row 0 and nothing outside the
    // domain has clean rows; a region whose columns hold no CPML column, no
    // column 0 and nothing outside the domain has clean columns.  Each wave
    // runs the copy of the step loop compiled for its case, the branch-free
    // one when both are clean; every copy meets the same barriers.  Only the
    // CPML terms and sources use cE4 / cH4, so clean waves load them only
    // at source cells.
    const bool rows_clean = gi0 + r0 >= P + 1 && gi0 + r0 + CPT <= NI - P;   // wave-uniform
    const bool cols_clean = gj0 >= P + 1 && gj0 + 64 <= NJ - P;              // region-uniform
    const bool need4 = !(rows_clean && cols_clean);

    float ps[4][CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int r = r0 + m, gi = gi0 + r;
        const bool ex = colx && gi >= 0 && gi <= NI;
        const size_t g = shot_off + (size_t)gi * pitch + gj;
        ez[m] = ex ? gmem(a.Ez_in)[g] : 0.f;
        hx[m] = ex ? gmem(a.Hx_in)[g] : 0.f;
        hy[m] = ex ? gmem(a.Hy_in)[g] : 0.f;
        const size_t c = (size_t)gi * pitch + gj;
        const auto* const cf = gmem(d.coef);
        cE0[m] = ex ? cf[c] : 0.f;
        cE1[m] = ex ? cf[pl + c] : 0.f;
        cE4[m] = (ex && need4) ? cf[2 * pl + c] : 0.f;
        cH1[m] = ex ? cf[3 * pl + c] : 0.f;
        cH4[m] = (ex && need4) ? cf[4 * pl + c] : 0.f;
        const bool ac = colact && gi >= 0 && gi < NI;
        if (ac) act |= 1u << m;
        const int qh = (ac && P > 0) ? xq_h(gi) : -1;
        const int qe = (ac && P > 0) ? xq_e(gi) : -1;
        ps[0][m] = qh >= 0 ? pin[(size_t)(gi < P ? HX0 : HXM) * xs + ((size_t)s * P + qh) * NJ + gj] : 0.f;
        ps[1][m] = qe >= 0 ? pin[(size_t)(gi <= P ? EX0 : EXM) * xs + ((size_t)s * P + qe) * NJ + gj] : 0.f;
        ps[2][m] = (ac && yqh >= 0) ? pin[4 * xs + (size_t)(yh_lo ? 2 : 3) * ys + ((size_t)s * NI + gi) * P + yqh] : 0.f;
        ps[3][m] = (ac && yqe >= 0) ? pin[4 * xs + (size_t)(ye_lo ? 0 : 1) * ys + ((size_t)s * NI + gi) * P + yqe] : 0.f;
    }
    const int cap = a.rx_list ? a.list_cap : 0;
    const auto* const blk = gmem(a.rx_list + (size_t)(s * ntiles + tile) * cap);
    const unsigned e_first = cap > 0 ? blk[lane] : kNoEntry;
    // CPML coefficient table entry of this thread (threads [0, 2P)), stored
    // to LDS once every load is in
    float4 xcv = make_float4(0.f, 0.f, 0.f, 0.f);
    if ((int)threadIdx.x < 2 * P) {
        const bool h = (int)threadIdx.x < P;
        const auto* b = gmem(d.pml + (h ? 4 * P : -P) + (int)threadIdx.x);
        xcv = make_float4(b[0] - 1.0f, b[P], b[2 * P], b[3 * P]);
    }

    // This wave's receivers (forward: tile-centre ones; adjoint: region
    // ones), compacted in list order with a ballot per 64 entries; the block
    // is packed, so a chunk with an empty entry is its last.
    int wn = 0;
    for (int base = 0; base < cap; base += 64) {
        const unsigned e = base == 0 ? e_first : blk[base + lane];
        const bool live = e != kNoEntry;
        const int li = (int)(e >> 26);
        const int ent = (live && li >= r0 && li < r0 + CPT) ? (((li - r0) << 26) | (int)(e & 0x3FFFFFFu)) : -1;
        const unsigned long long bal = __ballot(ent >= 0);
        const int pos = wn + (int)__popcll(bal & ((1ull << lane) - 1ull));
        if (ent >= 0 && pos < kWaveList) sRx[w][pos] = ent;
        wn += (int)__popcll(bal);
        if (__ballot(live) != ~0ull) break;
    }
    wn = wn < kWaveList ? wn : kWaveList;

    // CPML coefficients (RA - 1, RB, RE, RF) by slab index, H at [0, P), E at
    // [kPmax, kPmax + P), for both slab directions: one 16-byte LDS read per
    // CPML term (a broadcast for the row-uniform x-slab terms).
    __shared__ float4 sXC[2 * kPmax];
    if ((int)threadIdx.x < 2 * P)
        sXC[(int)threadIdx.x < P ? (int)threadIdx.x : kPmax - P + (int)threadIdx.x] = xcv;
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int o = (r0 + m) * 64 + lane;
#pragma unroll
        for (int c = 0; c < 4; ++c) sPsi[c][o] = ps[c][m];
    }
    sEzTop[w * 64 + lane] = ez[0];
    sHyBot[w * 64 + lane] = hy[CPT - 1];

    // one entry per lane, kWaveList / 64 registers
    int rxent[kWaveList / 64];
#pragma unroll
    for (int c = 0; c < kWaveList / 64; ++c) rxent[c] = (c * 64 + lane < wn) ? sRx[w][c * 64 + lane] : 0;

    // Adjoint: this wave's residual samples for every step of the launch,
    // loaded with the set-up into one LDS row of kWaveList floats per step,
    // so that the injection never waits on a global load inside the steps.
    __shared__ float sRes[ADJ ? W * K * kWaveList : 1];
    if constexpr (ADJ) {
#pragma unroll
        for (int c = 0; c < kWaveList / 64; ++c) {
            if (c * 64 + lane < wn) {
                const auto* rp = gmem(d.resid + (size_t)s * d.nt * d.nrx + (rxent[c] & 0xFFFFF));
                float rv[K];
#pragma unroll
                for (int k = 0; k < K; ++k) rv[k] = k < nsteps ? rp[(size_t)(n0 - k) * d.nrx] : 0.f;
#pragma unroll
                for (int k = 0; k < K; ++k) sRes[(w * K + k) * kWaveList + c * 64 + lane] = rv[k];
            }
        }
    }

    // Cells of this thread holding a source (forward).
    unsigned srcbits = 0u;
    if (!ADJ) {
        for (int kk = 0; kk < d.nsrc; ++kk) {
            const auto* sc = gmem(d.src + 2 * ((size_t)s * d.nsrc + kk));
            const int li = sc[0] - gi0 - r0;
            if (sc[1] == gj && li >= 0 && li < CPT) srcbits |= 1u << li;
        }
    }
    const bool src_any = __ballot(srcbits != 0u) != 0ull;   // wave-uniform
    if (!ADJ && src_any && !need4) {
#pragma unroll
        for (int m = 0; m < CPT; ++m)
            if (srcbits & (1u << m)) cE4[m] = gmem(d.coef)[2 * pl + (size_t)(gi0 + r0 + m) * pitch + gj];
    }

    const float idx = d.idx;
    double ae[CPT], as_[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) { ae[m] = 0.0; as_[m] = 0.0; }
    __syncthreads();

    // Row-uniform facts of each owned row, in scalar registers: bit m of rfl
    // = (gi >= 1), bit 16 + m = (gi <= P, the minus side); x-slab indices for
    // H and E (-1 outside the slabs and outside the domain).
    unsigned rfl = 0u;
    int qh_m[CPT], qe_m[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int gi = gi0 + r0 + m;
        const bool in = P > 0 && gi >= 0 && gi < NI;
        rfl |= (gi >= 1 ? 1u : 0u) << m;
        rfl |= (gi <= P ? 1u : 0u) << (16 + m);
        qh_m[m] = __builtin_amdgcn_readfirstlane(in ? xq_h(gi) : -1);
        qe_m[m] = __builtin_amdgcn_readfirstlane(in ? xq_e(gi) : -1);
    }
    rfl = __builtin_amdgcn_readfirstlane(rfl);
    auto rge1 = [&](int m) -> bool { return (rfl >> m) & 1u; };
    auto xminus = [&](int m) -> bool { return (rfl >> (16 + m)) & 1u; };

    // General H half-step.  Cells outside the domain (and the padding row and
    // column) have zero coefficients and zero fields, so the plain update
    // leaves them at zero; only the reference's row-0 / column-0 exceptions
    // and the CPML terms need conditions (rows: uniform per wave; columns:
    // per lane).
    //
    // XS: this wave's rows may hold row 0 or x-slab rows (row conditions and
    // x-slab terms needed); YS: the region's columns may hold column 0 or
    // y-slab columns.  Code for a side that cannot occur is compiled out.
    auto h_half = [&](auto xs_tag, auto ys_tag) {
        constexpr bool XS = decltype(xs_tag)::value, YS = decltype(ys_tag)::value;
        const float ez_below = (w + 1 < W) ? sEzTop[(w + 1) * 64 + lane] : 0.f;
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const float e = ez[m];
            const float dj = lane_next(e) - e;
            const float di = ((m + 1 < CPT) ? ez[m + 1 < CPT ? m + 1 : m] : ez_below) - e;
            const int o = (r0 + m) * 64 + lane;
            float vx = (!XS || rge1(m)) ? upd_hx(hx[m], cH1[m], dj) : hx[m];
            float vy = (!YS || gj >= 1) ? upd_hy(hy[m], cH1[m], di) : hy[m];
            if (YS && yqh >= 0) {
                const float4 yc = sXC[yqh];
                const float dd = dj * idx;
                const float ph = sPsi[2][o];
                vx = fmaf(-cH4[m], cpml_t(yc.x, yc.y, dd, ph), vx);
                sPsi[2][o] = cpml_psi(yc.z, yc.w, dd, ph);
            }
            const int q = XS ? qh_m[m] : -1;
            if (XS && q >= 0) {
                const float4 xc = sXC[q];
                const float dd = di * idx;
                const float ph = sPsi[0][o];
                vy = fmaf(cH4[m], cpml_t(xc.x, xc.y, dd, ph), vy);
                sPsi[0][o] = cpml_psi(xc.z, xc.w, dd, ph);
            }
            hx[m] = vx;
            hy[m] = vy;
        }
        sHyBot[w * 64 + lane] = hy[CPT - 1];
    };

    // General E update of owned cell m: plain update, the row-0 / column-0
    // exceptions, then the CPML terms in slab order x0, y0, xmax, ymax.
    auto e_cell = [&](auto xs_tag, auto ys_tag, int m, int o, float hyim, float hxjm) -> float {
        constexpr bool XS = decltype(xs_tag)::value, YS = decltype(ys_tag)::value;
        const float e0 = ez[m];
        const float dhy = hy[m] - hyim, dhx = hx[m] - hxjm;
        float v = upd_e(cE0[m], cE1[m], e0, dhy, dhx);
        if (XS && !rge1(m)) v = e0;
        if (YS) v = (gj >= 1) ? v : e0;
        const float c4 = cE4[m];
        auto yterm = [&]() {
            const float4 yc = sXC[kPmax + yqe];
            const float dyd = dhx * idx;
            const float p = sPsi[3][o];
            v = fmaf(-c4, cpml_t(yc.x, yc.y, dyd, p), v);
            sPsi[3][o] = cpml_psi(yc.z, yc.w, dyd, p);
        };
        const int q = XS ? qe_m[m] : -1;
        if (!XS || q < 0) {
            if (YS && yqe >= 0) yterm();
        } else {
            const float dxd = dhy * idx;
            auto xterm = [&]() {
                const float4 xc = sXC[kPmax + q];
                const float p = sPsi[1][o];
                v = fmaf(c4, cpml_t(xc.x, xc.y, dxd, p), v);
                sPsi[1][o] = cpml_psi(xc.z, xc.w, dxd, p);
            };
            if (xminus(m)) {
                xterm();
                if (YS && yqe >= 0) yterm();
            } else {
                if (YS && yqe >= 0 && ye_lo) yterm();
                xterm();
                if (YS && yqe >= 0 && !ye_lo) yterm();
            }
        }
        return v;
    };

    auto h_fast = [&]() {
        const float ez_below = (w + 1 < W) ? sEzTop[(w + 1) * 64 + lane] : 0.f;
#pragma unroll
        for (int m = 0; m < CPT; ++m) {
            const float e = ez[m];
            const float ezj = lane_next(e);
            const float ezi = (m + 1 < CPT) ? ez[m + 1 < CPT ? m + 1 : m] : ez_below;
            hx[m] = upd_hx(hx[m], cH1[m], ezj - e);
            hy[m] = upd_hy(hy[m], cH1[m], ezi - e);
        }
        sHyBot[w * 64 + lane] = hy[CPT - 1];
    };

    // Snapshot of (step n, this thread's row m) = snap_wave[slot(n) * snap_slot + m * pitch + gj]:
    // a wave-uniform base plus a per-lane offset; slots advance incrementally
    // (a ring of snap_mod slots, or n - snap_base), never by a division per step.
    const int64_t snap_slot = (int64_t)d.S * NI * pitch;
    float* const snap_wave = d.snaps ? d.snaps + ((int64_t)s * NI + gi0 + r0) * pitch : nullptr;
    // Ring size and base read once: in a multi-problem launch `d` lives in the
    // launch table in global memory, and a read inside the step loop would be
    // a vector load whose wait also drains the step's snapshot stores.
    const int snap_mod = d.snap_mod, snap_base = d.snap_base;
    auto slot_of = [&](int n) -> int { return snap_mod ? n % snap_mod : n - snap_base; };
    auto slot_next = [&](int sl) -> int { return (snap_mod && sl + 1 == snap_mod) ? 0 : sl + 1; };
    auto slot_prev = [&](int sl) -> int { return (snap_mod && sl == 0) ? snap_mod - 1 : sl - 1; };
    auto snap_ref = [&](int sl, int m) -> __attribute__((address_space(1))) float& {
        return gmem(snap_wave + (int64_t)sl * snap_slot + m * pitch)[gj];
    };
    // Time-blocked layout: per (slot, shot, tile) a block holding the waves
    // that own tile-centre rows and the tile-centre lanes only, ordered
    // [wave][centre lane][row], so each thread's CPT rows are contiguous (one
    // vector store / load per step) and each block is written by one
    // workgroup.  Halo lanes and halo-only waves take no memory.
    constexpr int kBlk = TBSnapGeom<RI, NT, K>::kBlock;
    constexpr int kWlo = TBSnapGeom<RI, NT, K>::kWaveLo;
    const bool tiled = d.snap_tiled != 0;
    const int64_t tslot = (int64_t)d.S * ntiles * kBlk;
    const bool wave_centre = r0 + CPT > K && r0 < K + TI;          // wave-uniform
    float* const tsnap = (d.snaps && wave_centre && centre_col)
                             ? d.snaps + ((int64_t)s * ntiles + tile) * kBlk +
                                   ((w - kWlo) * TJ + (lane - K)) * CPT
                             : nullptr;
    auto tstore = [&](int sl) {
        auto* const q = gmem(tsnap + sl * tslot);
#pragma unroll
        for (int m = 0; m < CPT; m += 4) {
            if constexpr (CPT % 4 == 0)
                *(__attribute__((address_space(1))) fvec4*)(q + m) =
                    fvec4{ez[m], ez[m + 1], ez[m + 2], ez[m + 3]};
        }
        if constexpr (CPT % 4 != 0)
#pragma unroll
            for (int m = 0; m < CPT; ++m) q[m] = ez[m];
    };
    auto tload = [&](int sl, float* f) {
        const auto* const q = gmem(tsnap + sl * tslot);
#pragma unroll
        for (int m = 0; m < CPT; m += 4) {
            if constexpr (CPT % 4 == 0) {
                const fvec4 v = *(const __attribute__((address_space(1))) fvec4*)(q + m);
                f[m] = v.x; f[m + 1] = v.y; f[m + 2] = v.z; f[m + 3] = v.w;
            }
        }
        if constexpr (CPT % 4 != 0)
#pragma unroll
            for (int m = 0; m < CPT; ++m) f[m] = q[m];
    };

    // adjoint: snapshots of steps n (fcur) and n - 1 (fprv) of the current step
    float fcur[CPT], fprv[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) { fcur[m] = 0.f; fprv[m] = 0.f; }
    if (ADJ && nsteps > 0) {
        if (tiled) {
            if (wave_centre && centre_col) {
                tload(slot_of(n0), fcur);
                tload(slot_of(n0 - 1), fprv);
            }
        } else {
#pragma unroll
            for (int m = 0; m < CPT; ++m) {
                const int r = r0 + m;
                if ((act & (1u << m)) && r >= K && r < K + TI && centre_col) {
                    fcur[m] = snap_ref(slot_of(n0), m);
                    fprv[m] = snap_ref(slot_of(n0 - 1), m);
                }
            }
        }
    }

    auto steps = [&](auto xs_tag, auto ys_tag) {
        constexpr bool FAST = !decltype(xs_tag)::value && !decltype(ys_tag)::value;
        // forward: slot of step n; adjoint: slot of step n - 2 (the prefetch)
        int sl = ADJ ? (nsteps > 1 ? slot_of(n0 - 2) : 0) : (snap_wave ? slot_of(n0) : 0);
        for (int k = 0; k < nsteps; ++k) {
            const int n = ADJ ? n0 - k : n0 + k;
            if (!ADJ) {
                // receivers of this wave, sampled before the step
#pragma unroll
                for (int c = 0; c < kWaveList / 64; ++c) {
                    if (c * 64 >= wn) break;                 // wave-uniform: no receivers left
                    const int cnt = wn - c * 64 < 64 ? wn - c * 64 : 64;
                    for (int q = 0; q < cnt; ++q) {
                        const int ent = __builtin_amdgcn_readlane(rxent[c], q);
                        if (lane == ((ent >> 20) & 63)) {
                            const int me = ent >> 26;
                            float v = 0.f;
#pragma unroll
                            for (int m = 0; m < CPT; ++m) v = (m == me) ? ez[m] : v;
                            const TBArgs* ar = reload(ap);
                            auto* const tr = gmem(ld_arg(&ar->d.traces));
                            tr[((size_t)s * ld_arg(&ar->d.nt) + n) * ld_arg(&ar->d.nrx) + (ent & 0xFFFFF)] = v;
                        }
                    }
                }
                if (FAST) h_fast(); else h_half(xs_tag, ys_tag);
                __syncthreads();
                const float hy_above = (w > 0) ? sHyBot[(w - 1) * 64 + lane] : 0.f;
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    const float hxjm = lane_prev(hx[m]);
                    const int r = r0 + m, gi = gi0 + r;
                    float v;
                    if (FAST) {
                        const float hyim = (m > 0) ? hy[m > 0 ? m - 1 : 0] : hy_above;
                        v = upd_e(cE0[m], cE1[m], ez[m], hy[m] - hyim, hx[m] - hxjm);
                    } else {
                        const float hyim = (m > 0) ? hy[m > 0 ? m - 1 : 0] : hy_above;
                        v = e_cell(xs_tag, ys_tag, m, r * 64 + lane, hyim, hxjm);
                    }
                    if (src_any && (srcbits & (1u << m))) {
                        const TBArgs* ar = reload(ap);
                        const int ns = ld_arg(&ar->d.nsrc), snt = ld_arg(&ar->d.nt);
                        const auto* const src = gmem(ld_arg(&ar->d.src));
                        const auto* const ssc = gmem(ld_arg(&ar->d.srcscale));
                        for (int kk = 0; kk < ns; ++kk) {
                            const auto* sc = src + 2 * ((size_t)s * ns + kk);
                            if (sc[0] == gi && sc[1] == gj) v = fmaf(-cE4[m], ssc[(size_t)kk * snt + n], v);
                        }
                    }
                    ez[m] = v;
                    if (!tiled && snap_wave && (FAST || (act & (1u << m))) && r >= K && r < K + TI &&
                        centre_col)
                        snap_ref(sl, m) = v;
                }
                if (tiled && tsnap && wave_centre && centre_col) tstore(sl);
                sEzTop[w * 64 + lane] = ez[0];
                sl = slot_next(sl);
                __syncthreads();
            } else {
                // adjoint source (back.cu:24-37) at receivers inside the region,
                // applied by the owner of the cell in list order
#pragma unroll
                for (int c = 0; c < kWaveList / 64; ++c) {
                    if (c * 64 >= wn) break;                 // wave-uniform: no receivers left
                    const int cnt = wn - c * 64 < 64 ? wn - c * 64 : 64;
                    const int mine = rxent[c];
                    const float gv = lane < cnt ? sRes[(w * K + k) * kWaveList + c * 64 + lane] : 0.f;
                    for (int q = 0; q < cnt; ++q) {
                        const int ent = __builtin_amdgcn_readlane(mine, q);
                        const float g = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, gv), q));
                        // x - 0 == x, so every row takes the subtraction and
                        // only the receiver's row changes (keeps ez[] in registers)
                        const int me = (lane == ((ent >> 20) & 63)) ? (ent >> 26) : -1;
#pragma unroll
                        for (int m = 0; m < CPT; ++m) ez[m] = ez[m] - ((m == me) ? g : 0.f);
                    }
                }
                const float hy_above = (w > 0) ? sHyBot[(w - 1) * 64 + lane] : 0.f;
#pragma unroll
                for (int m = 0; m < CPT; ++m) {
                    const float hxjm = lane_prev(hx[m]);
                    const int r = r0 + m;
                    float lam;
                    if (FAST) {
                        const float hyim = (m > 0) ? hy[m > 0 ? m - 1 : 0] : hy_above;
                        lam = upd_e(cE0[m], cE1[m], ez[m], hy[m] - hyim, hx[m] - hxjm);
                    } else {
                        const float hyim = (m > 0) ? hy[m > 0 ? m - 1 : 0] : hy_above;
                        lam = e_cell(xs_tag, ys_tag, m, r * 64 + lane, hyim, hxjm);
                    }
                    ez[m] = lam;
                    // Every row and lane accumulates (no select around the
                    // FP64 adds): only tile-centre cells are written back, and
                    // the other sums are never read.
                    ae[m] += (double)((fcur[m] - fprv[m]) * lam);
                    as_[m] += (double)(fcur[m] * lam);
                }
                // snapshots for the next step, loaded one step ahead
#pragma unroll
                for (int m = 0; m < CPT; ++m) fcur[m] = fprv[m];
                if (k + 1 < nsteps) {
                    if (tiled) {
                        if (wave_centre && centre_col) tload(sl, fprv);
                    } else {
#pragma unroll
                        for (int m = 0; m < CPT; ++m) {
                            const int r = r0 + m;
                            if ((FAST || (act & (1u << m))) && r >= K && r < K + TI && centre_col)
                                fprv[m] = snap_ref(sl, m);
                        }
                    }
                }
                sl = slot_prev(sl);
                sEzTop[w * 64 + lane] = ez[0];
                __syncthreads();
                if (FAST) h_fast(); else h_half(xs_tag, ys_tag);
                __syncthreads();
            }
        }
    };
    using Yes = std::true_type;
    using No = std::false_type;
    if (rows_clean) {
        if (cols_clean) steps(No{}, No{});
        else steps(No{}, Yes{});
    } else {
        if (cols_clean) steps(Yes{}, No{});
        else steps(Yes{}, Yes{});
    }

    // ---- write back the tile centre ----
    // output pointers and the sizes the write-back indexes with, read here:
    // nothing of the write-back is held in scalar registers through the steps
    const TBArgs* ar = reload(ap);
    auto* const pout = gmem(ld_arg(&ar->psi_out));
    auto* const Ez_out = gmem(ld_arg(&ar->Ez_out));
    auto* const Hx_out = gmem(ld_arg(&ar->Hx_out));
    auto* const Hy_out = gmem(ld_arg(&ar->Hy_out));
    auto* const gshot = gmem(ADJ ? ld_arg(&ar->gshot) : nullptr);
    const int wNI = ld_arg(&ar->d.NI), wNJ = ld_arg(&ar->d.NJ), wP = ld_arg(&ar->d.P);
    const int wS = ld_arg(&ar->d.S), wpitch = ld_arg(&ar->d.pitch);
    const size_t wxs = (size_t)wS * wP * wNJ, wys = (size_t)wS * wNI * wP;
    const size_t wshot = (size_t)s * (size_t)(wNI + 1) * wpitch;
#pragma unroll
    for (int m = 0; m < CPT; ++m) {
        const int r = r0 + m, gi = gi0 + r;
        if (!(r >= K && r < K + TI && centre_col)) continue;
        if (!(colx && gi >= 0 && gi <= wNI)) continue;
        const size_t g = wshot + (size_t)gi * wpitch + gj;
        Ez_out[g] = ez[m];
        Hx_out[g] = hx[m];
        Hy_out[g] = hy[m];
        if (!(act & (1u << m))) continue;
        const int o = r * 64 + lane;
        // x-slab indices as in the set-up (xq_h, xq_e)
        const int qh = wP > 0 ? (gi < wP ? wP - 1 - gi : (gi >= wNI - wP ? gi - (wNI - wP) : -1)) : -1;
        const int qe = wP > 0 ? ((gi >= 1 && gi <= wP) ? wP - gi : (gi >= wNI - wP ? gi - (wNI - wP) : -1)) : -1;
        if (qh >= 0) pout[(size_t)(gi < wP ? HX0 : HXM) * wxs + ((size_t)s * wP + qh) * wNJ + gj] = sPsi[0][o];
        if (qe >= 0) pout[(size_t)(gi <= wP ? EX0 : EXM) * wxs + ((size_t)s * wP + qe) * wNJ + gj] = sPsi[1][o];
        if (yqh >= 0) pout[4 * wxs + (size_t)(yh_lo ? 2 : 3) * wys + ((size_t)s * wNI + gi) * wP + yqh] = sPsi[2][o];
        if (yqe >= 0) pout[4 * wxs + (size_t)(ye_lo ? 0 : 1) * wys + ((size_t)s * wNI + gi) * wP + yqe] = sPsi[3][o];
        if constexpr (ADJ) {
            // Each cell of a shot has one owner per launch (the tile whose
            // centre holds it) and launches are ordered, so the sum is the
            // same as a load-add-store; the add happens at the memory side
            // and the workgroup does not wait for a read.
            auto* const g0 = gshot + ((size_t)2 * s * wNI + gi) * wNJ + gj;
            __builtin_amdgcn_global_atomic_fadd_f64(g0, ae[m]);
            __builtin_amdgcn_global_atomic_fadd_f64(g0 + (size_t)wNI * wNJ, as_[m]);
        }
    }
}

template <int RI, int NT, int K, bool ADJ>
__global__ void __launch_bounds__(NT)
k_tb(TBArgs a, int koff, int dir, int nsteps) {
    // `a` is the first kernel argument: at offset 0 of the argument segment
    const TBArgs* ap = (const TBArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    tb_body<RI, NT, K, ADJ>(a, ap, step_index(a.d, koff, dir), nsteps, blockIdx.x,
                            a.ti_count * a.tj_count * a.d.S);
}

// One launch over many problems (different grids, shots and time steps):
// the table holds the workgroup prefix sums and one entry per problem.
struct MultiEntry {
    TBArgs a;
    int n0;
    int nsteps;
    int pad[2];
};

__host__ __device__ inline size_t tb_table_header(int nitems) {
    return (((size_t)(nitems + 1) * sizeof(int)) + 63) / 64 * 64;
}

template <int RI, int NT, int K, bool ADJ>
__global__ void __launch_bounds__(NT)
k_tb_multi(const unsigned char* __restrict__ table, int nitems) {
    const int* prefix = (const int*)table;
    const int b = blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const MultiEntry* ents = (const MultiEntry*)(table + tb_table_header(nitems));
    const MultiEntry& me = ents[lo];
    tb_body<RI, NT, K, ADJ>(me.a, &me.a, me.n0, me.nsteps, b - prefix[lo], prefix[lo + 1] - prefix[lo]);
}

// Row-major Ez (S x (NI+1) x pitch) -> one slot of the time-blocked
// snapshot layout of variant (RI, NT, K): seeds a checkpoint segment's ring
// slot from a saved state.  Cells outside the domain are written as zero.
__global__ void __launch_bounds__(256) k_snap_pack(Dev d, int RI, int NT, int K, int ti_count,
                                                   int tj_count, int slot, const float* __restrict__ Ez) {
    const int ntiles = ti_count * tj_count, cpt = RI / (NT / 64);
    const int TJ = kTBCols - 2 * K - 1, wlo = K / cpt;
    const int blk = (int)tb_snap_block(RI, NT, K);
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)d.S * ntiles * blk) return;
    const int64_t b = e / blk;
    const int rem = (int)(e - b * blk);
    const int wl = rem / cpt, m = rem - wl * cpt;
    const int w = wlo + wl / TJ, lane = K + wl % TJ;
    const int sh = (int)(b / ntiles), tile = (int)(b - (int64_t)sh * ntiles);
    const int ti = tile / tj_count, tj = tile - ti * tj_count;
    const int gi = ti * (RI - 2 * K - 1) - K + w * cpt + m;
    const int gj = tj * TJ - K + lane;
    const bool in = gi >= 0 && gi < d.NI && gj >= 0 && gj < d.NJ;
    d.snaps[(int64_t)slot * d.S * ntiles * blk + e] =
        in ? Ez[((int64_t)sh * (d.NI + 1) + gi) * d.pitch + gj] : 0.f;
}

// Sequential FP32 sums of n strided elements of a and b (the CPML sigma_max
// means, compute.py:278-281), one thread, in index order.
__global__ void k_row_sums(const float* __restrict__ a, const float* __restrict__ b, int n, int stride,
                           float* __restrict__ out) {
    if (threadIdx.x != 0) return;
    float sa = 0.f, sb = 0.f;
    for (int j = 0; j < n; ++j) {
        sa = sa + a[(int64_t)j * stride];
        sb = sb + b[(int64_t)j * stride];
    }
    out[0] = sa;
    out[1] = sb;
}

// grad[c] += sum_s gshot[s][p][c] in shot order (deterministic).
__global__ void k_reduce_shots(const double* __restrict__ gshot, int S, int64_t ncell,
                               double* grad_eps, double* grad_sig) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    double se = 0.0, ss = 0.0;
    for (int s = 0; s < S; ++s) {
        se += gshot[(size_t)2 * s * ncell + c];
        ss += gshot[((size_t)2 * s + 1) * ncell + c];
    }
    grad_eps[c] += se;
    grad_sig[c] += ss;
}

// ---------------------------------------------------------------------------
// Launch helpers.
dim3 cell_grid(const fgpr_problem& p) {
    return dim3((p.NJ + kBlockJ - 1) / kBlockJ, (p.NI + kBlockI - 1) / kBlockI, 1);
}

int launch_forward_step(const Dev& d, const fgpr_problem& p, int koff, int dir, hipStream_t st) {
    const dim3 grid = cell_grid(p), blk(kBlockJ, kBlockI, 1);
    hipLaunchKernelGGL(k_h_phase<true>, grid, blk, 0, st, d, koff, dir);
    if (p.snaps)
        hipLaunchKernelGGL((k_e_phase<true, true, false>), grid, blk, 0, st, d, koff, dir);
    else
        hipLaunchKernelGGL((k_e_phase<true, false, false>), grid, blk, 0, st, d, koff, dir);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int launch_adjoint_step(const Dev& d, const fgpr_problem& p, int koff, int dir, hipStream_t st) {
    const dim3 grid = cell_grid(p), blk(kBlockJ, kBlockI, 1);
    const int ninj = p.S * p.nrx;
    if (ninj > 0)
        hipLaunchKernelGGL(k_adj_inject, dim3((ninj + 255) / 256), dim3(256), 0, st, d, koff, dir);
    hipLaunchKernelGGL((k_e_phase<false, false, true>), grid, blk, 0, st, d, koff, dir);
    hipLaunchKernelGGL(k_h_phase<false>, grid, blk, 0, st, d, koff, dir);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int launch_exact_step(const Dev& d, const fgpr_problem& p, int koff, int dir, hipStream_t st) {
    const dim3 grid = cell_grid(p), blk(kBlockJ, kBlockI, 1);
    hipLaunchKernelGGL(k_xadj_a, grid, blk, 0, st, d, koff, dir, p.scratch);
    hipLaunchKernelGGL(k_xadj_b, grid, blk, 0, st, d, p.scratch);
    hipLaunchKernelGGL(k_xadj_c, grid, blk, 0, st, d, p.scratch);
    const int ninj = p.S * p.nrx;
    if (ninj > 0)
        hipLaunchKernelGGL(k_xadj_record, dim3((ninj + 255) / 256), dim3(256), 0, st, d, koff, dir);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

// ---- time-blocked launches ----
using TBKernel = void (*)(TBArgs, int, int, int);

#define FGPR_TB_CASE(RI_, NT_, K_)                                                  \
    if (ri == RI_ && nt == NT_ && k == K_)                                         \
        return adj ? (TBKernel)k_tb<RI_, NT_, K_, true> : (TBKernel)k_tb<RI_, NT_, K_, false>;

TBKernel tb_kernel(int ri, int nt, int k, bool adj) {
    FGPR_TB_CASE(32, 256, 1) FGPR_TB_CASE(32, 256, 2) FGPR_TB_CASE(32, 256, 3)
    FGPR_TB_CASE(32, 256, 4) FGPR_TB_CASE(32, 256, 6)
    FGPR_TB_CASE(64, 512, 1) FGPR_TB_CASE(64, 512, 2) FGPR_TB_CASE(64, 512, 3)
    FGPR_TB_CASE(64, 512, 4) FGPR_TB_CASE(64, 512, 6) FGPR_TB_CASE(64, 512, 8)
    FGPR_TB_CASE(64, 256, 2) FGPR_TB_CASE(64, 256, 4) FGPR_TB_CASE(64, 256, 8)
    FGPR_TB_CASE(32, 512, 2) FGPR_TB_CASE(32, 512, 4) FGPR_TB_CASE(32, 512, 6)
    FGPR_TB_CASE(64, 1024, 2) FGPR_TB_CASE(64, 1024, 4) FGPR_TB_CASE(64, 1024, 6)
    FGPR_TB_CASE(64, 1024, 8)
    return nullptr;
}
#undef FGPR_TB_CASE

using TBMultiKernel = void (*)(const unsigned char*, int);

#define FGPR_TBM_CASE(RI_, NT_, K_)                                                 \
    if (ri == RI_ && nt == NT_ && k == K_)                                         \
        return adj ? (TBMultiKernel)k_tb_multi<RI_, NT_, K_, true>                 \
                   : (TBMultiKernel)k_tb_multi<RI_, NT_, K_, false>;

TBMultiKernel tb_multi_kernel(int ri, int nt, int k, bool adj) {
    FGPR_TBM_CASE(32, 512, 2) FGPR_TBM_CASE(32, 512, 4) FGPR_TBM_CASE(32, 512, 6)
    FGPR_TBM_CASE(64, 1024, 2) FGPR_TBM_CASE(64, 1024, 4) FGPR_TBM_CASE(64, 1024, 6)
    FGPR_TBM_CASE(64, 1024, 8)
    FGPR_TBM_CASE(64, 512, 2) FGPR_TBM_CASE(64, 512, 4) FGPR_TBM_CASE(64, 512, 6)
    FGPR_TBM_CASE(64, 512, 8)
    return nullptr;
}
#undef FGPR_TBM_CASE

bool tb_enabled(const fgpr_problem& p) { return p.tb_k > 0; }

// Steps per captured graph on the time-blocked path: whole pairs of launches
// (the ping-pong parity returns to 0), and enough launches that graph
// replay and the counter update stay small next to the kernels (16 launches).
int tb_chunk(int chunk, int k) {
    if (chunk <= 0) return chunk;
    const int pair = 2 * k;
    chunk = chunk < 16 * k ? 16 * k : chunk;
    return (chunk + pair - 1) / pair * pair;
}

int tb_check(const fgpr_problem& p, bool adj) {
    if (!tb_kernel(p.tb_ri, p.tb_nt, p.tb_k, adj))
        return fail(FGPR_EINVAL, "unsupported time-blocking variant (tb_ri, tb_nt, tb_k)");
    const int TI = p.tb_ri - 2 * p.tb_k - 1, TJ = kTBCols - 2 * p.tb_k - 1;
    if (p.tb_ti != (p.NI + 1 + TI - 1) / TI || p.tb_tj != (p.NJ + 1 + TJ - 1) / TJ)
        return fail(FGPR_EINVAL, "tb_ti / tb_tj do not match the tile size of the variant");
    if (!p.Ez2 || !p.Hx2 || !p.Hy2 || (p.P > 0 && !p.psi2))
        return fail(FGPR_EINVAL, "time-blocked path needs the ping-pong buffers");
    if (p.P > kPmax) return fail(FGPR_EINVAL, "CPML thicker than 128 cells");
    if (p.nrx >= (1 << 20)) return fail(FGPR_EINVAL, "time-blocked path holds at most 2^20 - 1 receivers per shot");
    if (adj && !p.gshot) return fail(FGPR_EINVAL, "time-blocked adjoint needs gshot");
    if ((adj ? !p.tb_inj_list : !p.tb_rx_list) && p.nrx > 0)
        return fail(FGPR_EINVAL, "time-blocked path needs the per-tile receiver blocks");
    if (p.nrx > 0 && p.tb_list_cap != 64 && p.tb_list_cap != 128)
        return fail(FGPR_EINVAL, "tb_list_cap must be 64 or 128");
    return FGPR_OK;
}

TBArgs make_tb(const fgpr_problem& p, const Dev& d, int parity, bool adj) {
    TBArgs a;
    a.d = d;
    const float* inE = parity ? p.Ez2 : p.Ez;
    const float* inX = parity ? p.Hx2 : p.Hx;
    const float* inY = parity ? p.Hy2 : p.Hy;
    const float* inP = parity ? p.psi2 : p.psi;
    a.Ez_in = inE; a.Hx_in = inX; a.Hy_in = inY; a.psi_in = inP;
    a.Ez_out = parity ? p.Ez : p.Ez2;
    a.Hx_out = parity ? p.Hx : p.Hx2;
    a.Hy_out = parity ? p.Hy : p.Hy2;
    a.psi_out = parity ? p.psi : p.psi2;
    a.ti_count = p.tb_ti;
    a.tj_count = p.tb_tj;
    // Interior tiles: gi0 = ti TI - K with gi0 >= P + 1 and gi0 + RI <= NI - P
    // (and the same for columns), a rectangle of tile indices.
    {
        const int TI = p.tb_ri - 2 * p.tb_k - 1, TJ = kTBCols - 2 * p.tb_k - 1;
        auto range = [](int lo_num, int hi_num, int T, int count, int& first, int& n) {
            first = (lo_num + T - 1) / T;                       // lo_num >= 1
            const int last = hi_num < 0 ? -1 : (hi_num / T < count - 1 ? hi_num / T : count - 1);
            n = last >= first ? last - first + 1 : 0;
        };
        range(p.P + 1 + p.tb_k, p.NI - p.P - p.tb_ri + p.tb_k, TI, p.tb_ti, a.int_ti0, a.int_ni);
        range(p.P + 1 + p.tb_k, p.NJ - p.P - kTBCols + p.tb_k, TJ, p.tb_tj, a.int_tj0, a.int_nj);
        if (a.int_ni == 0 || a.int_nj == 0) { a.int_ti0 = 0; a.int_ni = 0; a.int_tj0 = 0; a.int_nj = 0; }
    }
    a.rx_list = (const unsigned*)(p.nrx > 0 ? (adj ? p.tb_inj_list : p.tb_rx_list) : nullptr);
    a.list_cap = p.tb_list_cap;
    a.gshot = p.gshot;
    a.fd_S = make_fastdiv(p.S);
    a.fd_nj = make_fastdiv(a.int_nj);
    a.fd_per = make_fastdiv(a.tj_count - a.int_nj);
    a.fd_tj = make_fastdiv(a.tj_count);
    return a;
}

// After an odd number of ping-pong launches the state sits in the partner
// buffers; copy it back so every call leaves it in Ez/Hx/Hy/psi.
int tb_settle(const fgpr_problem& p, int parity, hipStream_t st) {
    if (!parity) return FGPR_OK;
    const size_t fb = (size_t)p.S * (p.NI + 1) * p.pitch * sizeof(float);
    HIP_TRY(hipMemcpyAsync(p.Ez, p.Ez2, fb, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(p.Hx, p.Hx2, fb, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(p.Hy, p.Hy2, fb, hipMemcpyDeviceToDevice, st));
    if (p.P > 0)
        HIP_TRY(hipMemcpyAsync(p.psi, p.psi2, (size_t)fgpr_psi_floats(p.NI, p.NJ, p.P, p.S) * sizeof(float),
                               hipMemcpyDeviceToDevice, st));
    return FGPR_OK;
}

int launch_tb(const fgpr_problem& p, const Dev& d, bool adj, int parity, int koff, int dir,
              int nsteps, hipStream_t st) {
    TBKernel kfn = tb_kernel(p.tb_ri, p.tb_nt, p.tb_k, adj);
    const TBArgs a = make_tb(p, d, parity, adj);
    const int nwg = p.tb_ti * p.tb_tj * p.S;
    hipLaunchKernelGGL(kfn, dim3(nwg), dim3(p.tb_nt), 0, st, a, koff, dir, nsteps);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

// Captured graphs: one per (descriptor, direction, chunk).  The key lists
// the descriptor's fields by name -- the device pointers baked into the
// graph and the sizes -- never its raw bytes (the struct has padding).  The
// cache holds at most g_graph_limit executables: past that, the least
// recently launched one is destroyed once its last launch has finished (an
// event recorded after every launch), so callers whose buffers move between
// calls do not accumulate graphs.
static_assert(sizeof(fgpr_problem) == 264, "a descriptor field was added: extend make_key");

struct GraphKey {
    std::vector<int64_t> v;
    bool operator<(const GraphKey& o) const { return v < o.v; }
};

struct GraphEntry {
    hipGraphExec_t ex = nullptr;
    hipEvent_t done = nullptr;
    uint64_t used = 0;
    bool evictable = true;
};

std::mutex g_graph_mu;
std::map<GraphKey, GraphEntry> g_graphs;
// Graphs of fgpr_profile_kernel's timed launches: a cache of their own, so
// timing (the autotuner) never evicts a sweep's graphs.
std::map<GraphKey, GraphEntry> g_prof_graphs;
constexpr size_t kProfGraphLimit = 8;
// Graphs of fgpr_tv's launch sequence (about 2 per ADMM iteration).
std::map<GraphKey, GraphEntry> g_tv_graphs;
constexpr size_t kTVGraphLimit = 8;
uint64_t g_graph_clock = 0;
size_t g_graph_limit = 64;

GraphKey make_key(const fgpr_problem& p, int kind, int chunk) {
    GraphKey k;
    auto I = [&](int64_t x) { k.v.push_back(x); };
    auto P = [&](const void* x) { k.v.push_back((int64_t)(uintptr_t)x); };
    auto F = [&](float x) { int32_t b; std::memcpy(&b, &x, 4); k.v.push_back(b); };
    I(p.NI); I(p.NJ); I(p.P); I(p.S); I(p.nt); I(p.nsrc); I(p.nrx); I(p.pitch);
    F(p.dx); F(p.dt);
    P(p.coef); P(p.pml); P(p.src); P(p.srcscale); P(p.rx); P(p.Ez); P(p.Hx); P(p.Hy); P(p.psi);
    P(p.traces); P(p.snaps); P(p.resid); P(p.grad_eps); P(p.grad_sig); P(p.ctr);
    P(p.Ez2); P(p.Hx2); P(p.Hy2); P(p.psi2);
    I(p.tb_ri); I(p.tb_nt); I(p.tb_k); I(p.tb_ti); I(p.tb_tj); I(p.tb_list_cap);
    P(p.tb_rx_list); P(p.tb_inj_list); P(p.gshot); P(p.scratch);
    I(p.snap_base); I(p.nsnap); I(p.snap_mod); I(p.snap_tiled);
    I(kind); I(chunk);
    return k;
}

void destroy_entry(GraphEntry& e) {
    if (e.done) {
        (void)hipEventSynchronize(e.done);
        (void)hipEventDestroy(e.done);
    }
    if (e.ex) (void)hipGraphExecDestroy(e.ex);
    e.ex = nullptr;
    e.done = nullptr;
}

// Least recently launched entries go first; `keep` is never evicted.
void evict_over_limit(std::map<GraphKey, GraphEntry>& cache, size_t limit, const GraphEntry* keep) {
    while (cache.size() > limit) {
        auto victim = cache.end();
        for (auto it = cache.begin(); it != cache.end(); ++it)
            if (&it->second != keep && it->second.evictable &&
                (victim == cache.end() || it->second.used < victim->second.used))
                victim = it;
        if (victim == cache.end()) return;
        destroy_entry(victim->second);
        cache.erase(victim);
    }
}

// Capture what `body` launches on a private stream into one executable
// graph.  Called with g_graph_mu held.
template <typename Body>
int capture_body(const Body& body, hipGraphExec_t* out) {
    hipStream_t cs;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(FGPR_EHIP, "hipStreamBeginCapture failed");
    }
    const int rc = body(cs);
    hipError_t ec = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != FGPR_OK) {
        if (ec == hipSuccess) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) return fail(FGPR_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipGraphExec_t ex;
    ec = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ec != hipSuccess) return fail(FGPR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
    *out = ex;
    return FGPR_OK;
}

// Launch the cached graph of `key` `reps` times on st, capturing `body` on
// first use; records the entry's completion event after the launches.
template <typename Body>
int launch_cached(std::map<GraphKey, GraphEntry>& cache, size_t limit, GraphKey key,
                  const Body& body, int reps, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    auto it = cache.find(key);
    if (it == cache.end()) {
        GraphEntry e;
        int rc = capture_body(body, &e.ex);
        if (rc) return rc;
        if (hipEventCreateWithFlags(&e.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGraphExecDestroy(e.ex);
            return fail(FGPR_EHIP, "hipEventCreateWithFlags failed");
        }
        it = cache.emplace(std::move(key), e).first;
    }
    GraphEntry& e = it->second;
    for (int r = 0; r < reps; ++r) HIP_TRY(hipGraphLaunch(e.ex, st));
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(st, &cap));
    if (cap == hipStreamCaptureStatusNone) {
        HIP_TRY(hipEventRecord(e.done, st));
    } else {
        e.evictable = false;          // launched into the caller's capture: kept until released
    }
    e.used = ++g_graph_clock;
    evict_over_limit(cache, limit, &e);
    return FGPR_OK;
}

// The sweep graphs: kind 0 = forward (counter advances by +chunk),
// 1 = reference adjoint, 2 = exact adjoint (both -chunk); `chunk` steps from
// the step the descriptor's counter holds.
int launch_graph(const fgpr_problem& p, int kind, int chunk, int reps, hipStream_t st) {
    auto body = [&](hipStream_t cs) {
        const Dev d = to_dev(p);
        const int dir = kind == 0 ? 1 : -1;
        int rc = FGPR_OK;
        if (kind == 2) {
            for (int k = 0; k < chunk && rc == FGPR_OK; ++k) rc = launch_exact_step(d, p, k, dir, cs);
        } else if (tb_enabled(p)) {
            // chunk is a multiple of 2*tb_k: an even number of ping-pong launches,
            // so every replay starts from the primary buffers.
            for (int c = 0; c * p.tb_k < chunk && rc == FGPR_OK; ++c)
                rc = launch_tb(p, d, kind == 1, c & 1, c * p.tb_k, dir, p.tb_k, cs);
        } else {
            for (int k = 0; k < chunk && rc == FGPR_OK; ++k)
                rc = kind == 0 ? launch_forward_step(d, p, k, dir, cs) : launch_adjoint_step(d, p, k, dir, cs);
        }
        if (rc == FGPR_OK) hipLaunchKernelGGL(k_advance_ctr, dim3(1), dim3(64), 0, cs, p.ctr, dir * chunk);
        return rc;
    };
    return launch_cached(g_graphs, g_graph_limit, make_key(p, kind, chunk), body, reps, st);
}

int check_problem(const fgpr_problem* p) {
    if (!p) return fail(FGPR_EINVAL, "null problem");
    if (p->NI < 2 || p->NJ < 2) return fail(FGPR_EINVAL, "grid must be at least 2x2");
    if (p->P < 0) return fail(FGPR_EINVAL, "negative CPML thickness");
    if (p->P > 0 && (2 * p->P >= p->NI || 2 * p->P >= p->NJ))
        return fail(FGPR_EINVAL, "CPML slabs overlap: need NI > 2P and NJ > 2P");
    if (p->S < 1 || p->nt < 1 || p->nsrc < 0 || p->nrx < 0)
        return fail(FGPR_EINVAL, "bad S / nt / nsrc / nrx");
    if (p->pitch != fgpr_pitch(p->NJ)) return fail(FGPR_EINVAL, "pitch must equal fgpr_pitch(NJ)");
    if (!p->coef || !p->Ez || !p->Hx || !p->Hy) return fail(FGPR_EINVAL, "null coef/field pointer");
    if (p->P > 0 && (!p->pml || !p->psi)) return fail(FGPR_EINVAL, "null pml/psi pointer");
    if (!(p->dx > 0.f)) return fail(FGPR_EINVAL, "dx must be positive");
    if (p->snap_tiled && p->tb_k <= 0)
        return fail(FGPR_EINVAL, "time-blocked snapshot layout (snap_tiled) needs the time-blocked path");
    return FGPR_OK;
}


// ---------------------------------------------------------------------------
// Total-variation denoiser (src/tv.py:135-217): ADMM on the max-normalised
// image, zero-padded by one cell, periodic differences on the padded grid,
// soft-thresholded split variables.  All iterations run on the device as a
// sequence of launches; the stopping test ||X - X_prev|| < tol is evaluated
// on the device (fixed-order FP64 sum of per-block partials), and every later
// launch returns at once.  The float32
// operation order follows the reference expressions; the 4-neighbour sum
// (conv2d in the reference, whose summation order is the library's) is
// (up + down) + (left + right).
struct TVWork {
    float* X[2];      // ping-pong iterates, (M+2) x (N+2)
    float* Y;         // normalised data, zero border
    float* Zx; float* Zy; float* Ux; float* Uy;
    float* scal;      // [0] data max
    int* flags;       // [0] done, [1] iteration that converged
    double* part;     // one partial sum of squares per X-sweep block
    int M2, N2;
};

constexpr int kTVBx = 64, kTVBy = 4;

__global__ void __launch_bounds__(1024) k_tv_max(const float* __restrict__ data, int64_t n, TVWork w) {
    __shared__ float red[1024];
    float m = -INFINITY;
    for (int64_t k = threadIdx.x; k < n; k += 1024) m = fmaxf(m, data[k]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        w.scal[0] = red[0];
        w.flags[0] = 0;
        w.flags[1] = -1;
    }
}

__global__ void __launch_bounds__(kTVBx * kTVBy) k_tv_init(const float* __restrict__ data, TVWork w) {
    const int j = blockIdx.x * kTVBx + threadIdx.x, i = blockIdx.y * kTVBy + threadIdx.y;
    if (i >= w.M2 || j >= w.N2) return;
    const size_t c = (size_t)i * w.N2 + j;
    const int N = w.N2 - 2;
    const bool in = i >= 1 && i < w.M2 - 1 && j >= 1 && j < w.N2 - 1;
    const float den = w.scal[0] + 1e-6f;
    const float y = in ? data[(size_t)(i - 1) * N + (j - 1)] / den : 0.f;
    w.X[0][c] = y; w.X[1][c] = y; w.Y[c] = y;
    w.Zx[c] = 0.f; w.Zy[c] = 0.f; w.Ux[c] = 0.f; w.Uy[c] = 0.f;
}

// One interior cell of the X sweep (Jacobi from Xo into Xn); returns the
// squared change.
__device__ __forceinline__ double tv_x_cell(const TVWork& w, const float* __restrict__ Xo,
                                            float* __restrict__ Xn, int i, int j, float lam,
                                            float lr, float den, float rho) {
    const int N2 = w.N2;
    const size_t c = (size_t)i * N2 + j;
    // Dxt Z = Z - Z[j+1], Dyt Z = Z - Z[i+1] (no wrap inside the interior)
    const float dxz = w.Zx[c] - w.Zx[c + 1], dyz = w.Zy[c] - w.Zy[c + N2];
    const float dxu = w.Ux[c] - w.Ux[c + 1], dyu = w.Uy[c] - w.Uy[c + N2];
    const float rhs = (w.Y[c] + lr * (dxz + dyz)) - lam * (dxu + dyu);
    const float nb = (Xo[c - N2] + Xo[c + N2]) + (Xo[c - 1] + Xo[c + 1]);
    const float x = ((nb * lam) * rho + rhs) / den;
    Xn[c] = x;
    const double e = (double)x - (double)Xo[c];
    return e * e;
}

// One cell of the split-variable and multiplier update from X (periodic
// backward differences over the padded grid).
__device__ __forceinline__ void tv_zu_cell(const TVWork& w, const float* __restrict__ X, int i,
                                           int j, float rho, float inv_rho) {
    const int M2 = w.M2, N2 = w.N2;
    const size_t c = (size_t)i * N2 + j;
    const float x = X[c];
    const float dxx = x - X[(size_t)i * N2 + (j > 0 ? j - 1 : N2 - 1)];
    const float dyx = x - X[(size_t)(i > 0 ? i - 1 : M2 - 1) * N2 + j];
    const float tx = (w.Ux[c] + rho * dxx) / rho;
    const float ty = (w.Uy[c] + rho * dyx) / rho;
    // sign(T) * clamp(|T| - 1/rho, min = 0)
    const float zx = (tx > 0.f ? 1.f : (tx < 0.f ? -1.f : 0.f)) * fmaxf(fabsf(tx) - inv_rho, 0.f);
    const float zy = (ty > 0.f ? 1.f : (ty < 0.f ? -1.f : 0.f)) * fmaxf(fabsf(ty) - inv_rho, 0.f);
    w.Zx[c] = zx;
    w.Zy[c] = zy;
    w.Ux[c] = w.Ux[c] + rho * (dxx - zx);
    w.Uy[c] = w.Uy[c] + rho * (dyx - zy);
}

// X sweep of iteration k: interior of X[(k+1)&1] from X[k&1], and per-block
// sums of (X_new - X_old)^2.
__global__ void __launch_bounds__(kTVBx * kTVBy) k_tv_x(TVWork w, int k, float lam, float lr,
                                                       float den, float rho) {
    if (w.flags[0]) return;
    const int j = blockIdx.x * kTVBx + threadIdx.x, i = blockIdx.y * kTVBy + threadIdx.y;
    double d2 = 0.0;
    if (i >= 1 && i < w.M2 - 1 && j >= 1 && j < w.N2 - 1)
        d2 = tv_x_cell(w, w.X[k & 1], w.X[(k + 1) & 1], i, j, lam, lr, den, rho);
    // block sum (fixed order); the next Z/U sweep adds the block sums in order
    __shared__ double red[kTVBx * kTVBy];
    const int t = threadIdx.y * kTVBx + threadIdx.x;
    red[t] = d2;
    __syncthreads();
    for (int s = kTVBx * kTVBy / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) w.part[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// Stopping test of iteration k: ||X_k - X_{k-1}|| < tol from the X sweep's
// block sums, added in fixed order by one workgroup (the first workgroup of
// the Z/U sweep that follows; a last-block pattern inside the X sweep would
// need device-scope fences, which write back L2 across the XCDs).
__device__ void tv_check(const TVWork& w, int k, int nparts, double tol) {
    __shared__ double red[kTVBx * kTVBy];
    const int t = threadIdx.y * kTVBx + threadIdx.x;
    double acc = 0.0;
    for (int b = t; b < nparts; b += kTVBx * kTVBy) acc += w.part[b];
    red[t] = acc;
    __syncthreads();
    for (int s = kTVBx * kTVBy / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0 && sqrt(red[0]) < tol) { w.flags[1] = k; w.flags[0] = 1; }
}

// Split-variable and multiplier update of iteration k from X[(k+1)&1]
// (periodic backward differences over the padded grid).
__global__ void __launch_bounds__(kTVBx * kTVBy) k_tv_zu(TVWork w, int k, float rho, float inv_rho,
                                                        int nparts, double tol) {
    if (w.flags[0]) return;
    if (blockIdx.x == 0 && blockIdx.y == 0) tv_check(w, k, nparts, tol);
    const int j = blockIdx.x * kTVBx + threadIdx.x, i = blockIdx.y * kTVBy + threadIdx.y;
    if (i >= w.M2 || j >= w.N2) return;
    tv_zu_cell(w, w.X[(k + 1) & 1], i, j, rho, inv_rho);
}

__global__ void __launch_bounds__(kTVBx * kTVBy) k_tv_out(TVWork w, int num_iter, float* __restrict__ out) {
    const int j = blockIdx.x * kTVBx + threadIdx.x, i = blockIdx.y * kTVBy + threadIdx.y;
    const int M = w.M2 - 2, N = w.N2 - 2;
    if (i >= M || j >= N) return;
    const int kl = w.flags[0] ? w.flags[1] : num_iter - 1;    // last iteration run
    const float* X = w.X[(kl + 1) & 1];
    out[(size_t)i * N + j] = X[(size_t)(i + 1) * w.N2 + (j + 1)] * w.scal[0];
}

// ---------------------------------------------------------------------------
// FIR low-pass of traces (multiscale.py:19-50): the reference reflect-pads
// each trace by L-1 samples on the left and correlates with the taps, i.e.
// out[t] = sum_k h[k] * x[|t + k - (L-1)|].  Traces are (S, nt, nrx) with the
// receiver fastest, so a thread per (s, t, r) reads coalesced rows; taps are
// summed in ascending order in FP32.
__global__ void __launch_bounds__(256) k_fir(const float* __restrict__ x, int S, int nt, int nrx,
                                             const float* __restrict__ h, int L,
                                             float* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)S * nt * nrx) return;
    const int r = (int)(e % nrx);
    const int64_t st = e / nrx;
    const int t = (int)(st % nt);
    const int64_t base = (st - t) * nrx + r;          // (s, 0, r)
    float acc = 0.f;
    for (int k = 0; k < L; ++k) {
        int i = t + k - (L - 1);
        i = i < 0 ? -i : i;
        acc = acc + h[k] * x[base + (int64_t)i * nrx];
    }
    out[e] = acc;
}

int64_t tv_blocks(int M, int N) {
    return (int64_t)((N + 2 + kTVBx - 1) / kTVBx) * ((M + 2 + kTVBy - 1) / kTVBy);
}

int64_t tv_bytes(int M, int N) {
    const int64_t plane = (int64_t)(M + 2) * (N + 2) * 4;
    const int64_t pl = (plane + 255) / 256 * 256;
    return 7 * pl + 256 + tv_blocks(M, N) * 8;
}

// Workspace of fgpr_tv_static: a file-scope device array (each device the
// library runs on holds its own copy until the process ends).  Host code
// never passes the array's name to a kernel: its device address, per
// device, comes from hipGetSymbolAddress on the first call on that device.
static __device__ __align__(256) unsigned char scratch[FGPR_TV_STATIC_BYTES];

std::mutex g_scratch_mu;
std::map<int, void*> g_scratch_ptr;

int scratch_address(void** out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it == g_scratch_ptr.end()) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(FGPR_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_ptr.emplace(dev, ptr).first;
    }
    *out = it->second;
    return FGPR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int fgpr_abi_version(void) { return FGPR_ABI_VERSION; }

size_t fgpr_problem_size(void) { return sizeof(fgpr_problem); }

const char* fgpr_last_error(void) { return g_err.c_str(); }

int fgpr_pitch(int NJ) { return ((NJ + 1 + 63) / 64) * 64; }

int64_t fgpr_psi_floats(int NI, int NJ, int P, int S) {
    return 4 * (int64_t)S * P * NJ + 4 * (int64_t)S * NI * P;
}

int fgpr_validate(const fgpr_problem* p) { return check_problem(p); }

int fgpr_tb_tiles(int NI, int NJ, int tb_ri, int tb_nt, int tb_k, int* TI, int* TJ,
                  int* tiles_i, int* tiles_j) {
    if (!tb_kernel(tb_ri, tb_nt, tb_k, false)) return fail(FGPR_EINVAL, "unsupported time-blocking variant");
    const int ti = tb_ri - 2 * tb_k - 1, tj = kTBCols - 2 * tb_k - 1;
    if (TI) *TI = ti;
    if (TJ) *TJ = tj;
    if (tiles_i) *tiles_i = (NI + 1 + ti - 1) / ti;
    if (tiles_j) *tiles_j = (NJ + 1 + tj - 1) / tj;
    return FGPR_OK;
}

int fgpr_cell_coeffs(const float* er, const float* se, const float* mr, int NI, int NJ,
                     float dt, float dx, int adjoint, float* coef, void* stream) {
    if (!er || !se || !mr || !coef) return fail(FGPR_EINVAL, "null pointer");
    if (NI < 1 || NJ < 1) return fail(FGPR_EINVAL, "bad grid");
    const int pitch = fgpr_pitch(NJ);
    hipLaunchKernelGGL(k_cell_coeffs, dim3((pitch + 63) / 64, NI + 1), dim3(64), 0,
                       (hipStream_t)stream, er, se, mr, NI, NJ, pitch, dt, dx, adjoint, coef);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_forward(const fgpr_problem* p, int n0, int n1, int mode, int chunk, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (n0 < 0 || n1 > p->nt || n0 > n1) return fail(FGPR_EINVAL, "bad step range");
    if (!p->traces && p->nrx > 0) return fail(FGPR_EINVAL, "null traces");
    if (p->nsrc > 0 && (!p->src || !p->srcscale)) return fail(FGPR_EINVAL, "null source arrays");
    if (p->snaps && p->snap_mod > 0 && p->nsnap < p->snap_mod)
        return fail(FGPR_EINVAL, "snapshot ring larger than the buffer (snap_mod > nsnap)");
    if (p->snaps && p->snap_mod == 0 &&
        (n0 - p->snap_base < 0 || (n1 > n0 && n1 - 1 - p->snap_base >= p->nsnap)))
        return fail(FGPR_EINVAL, "forward steps outside the snapshot buffer (snap_base, nsnap)");
    const bool tb = tb_enabled(*p);
    if (tb && (rc = tb_check(*p, false))) return rc;
    hipStream_t st = (hipStream_t)stream;
    fgpr_problem q = *p;
    int n = n0;
    if (tb) chunk = tb_chunk(chunk, p->tb_k);
    if (mode == FGPR_EXEC_GRAPH && chunk > 0 && p->ctr && n1 - n0 >= chunk) {
        const int reps = (n1 - n0) / chunk;
        hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(64), 0, st, q.ctr, n0);
        if ((rc = launch_graph(q, 0, chunk, reps, st))) return rc;
        n += reps * chunk;
    }
    q.ctr = nullptr;
    const Dev d = to_dev(q);
    if (tb) {
        int parity = 0;
        while (n < n1) {
            const int ns = (n1 - n < p->tb_k) ? n1 - n : p->tb_k;
            rc = launch_tb(q, d, false, parity, n, 1, ns, st);
            if (rc) return rc;
            parity ^= 1;
            n += ns;
        }
        return tb_settle(q, parity, st);
    }
    for (; n < n1; ++n) {
        rc = launch_forward_step(d, q, n, 1, st);
        if (rc) return rc;
    }
    return FGPR_OK;
}

int fgpr_adjoint(const fgpr_problem* p, int n_hi, int n_lo, int grad_mode, int mode, int chunk,
                 void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (grad_mode == FGPR_GRAD_EXACT && p->snap_tiled)
        return fail(FGPR_EINVAL, "the exact adjoint reads row-major snapshots (snap_tiled must be 0)");
    if (grad_mode != FGPR_GRAD_REFERENCE && grad_mode != FGPR_GRAD_EXACT)
        return fail(FGPR_EINVAL, "unknown gradient mode");
    const bool exact = grad_mode == FGPR_GRAD_EXACT;
    if (n_lo < (exact ? 0 : 1) || n_hi >= p->nt || n_hi < n_lo - 1) return fail(FGPR_EINVAL, "bad step range");
    // The time-blocked reference adjoint may leave the shot reduction to the
    // caller (both gradient pointers null): several shot groups of one batch
    // then run on separate streams and fgpr_reduce_shots sums them in order.
    const bool defer = !exact && tb_enabled(*p) && !p->grad_eps && !p->grad_sig;
    if (!p->snaps || (!defer && (!p->grad_eps || !p->grad_sig)))
        return fail(FGPR_EINVAL, "adjoint needs snaps and gradients");
    if (p->nrx > 0 && (!p->resid || !p->rx)) return fail(FGPR_EINVAL, "null residual/receivers");
    if (p->snap_mod > 0) {
        if (p->nsnap < p->snap_mod || n_hi - (n_lo >= 1 ? n_lo - 1 : 0) + 1 > p->snap_mod)
            return fail(FGPR_EINVAL, "adjoint range longer than the snapshot ring (snap_mod)");
    } else {
        const int lo = (n_lo >= 1 ? n_lo - 1 : 0) - p->snap_base, hi = n_hi - p->snap_base;
        if (n_hi >= n_lo && (lo < 0 || hi >= p->nsnap))
            return fail(FGPR_EINVAL, "adjoint steps outside the snapshot buffer (snap_base, nsnap)");
    }
    if (exact) {
        if (!p->scratch) return fail(FGPR_EINVAL, "exact adjoint needs the scratch planes");
        hipStream_t st = (hipStream_t)stream;
        fgpr_problem q = *p;
        int n = n_hi;
        if (mode == FGPR_EXEC_GRAPH && chunk > 0 && p->ctr && n_hi - n_lo + 1 >= chunk) {
            const int reps = (n_hi - n_lo + 1) / chunk;
            hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(64), 0, st, q.ctr, n_hi);
            if ((rc = launch_graph(q, 2, chunk, reps, st))) return rc;
            n -= reps * chunk;
        }
        q.ctr = nullptr;
        const Dev d = to_dev(q);
        for (; n >= n_lo; --n)
            if ((rc = launch_exact_step(d, q, n, 1, st))) return rc;
        return FGPR_OK;
    }
    const bool tb = tb_enabled(*p);
    if (tb && (rc = tb_check(*p, true))) return rc;
    hipStream_t st = (hipStream_t)stream;
    fgpr_problem q = *p;
    int n = n_hi;
    if (tb) chunk = tb_chunk(chunk, p->tb_k);
    if (mode == FGPR_EXEC_GRAPH && chunk > 0 && p->ctr && n_hi - n_lo + 1 >= chunk) {
        const int reps = (n_hi - n_lo + 1) / chunk;
        hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(64), 0, st, q.ctr, n_hi);
        if ((rc = launch_graph(q, 1, chunk, reps, st))) return rc;
        n -= reps * chunk;
    }
    q.ctr = nullptr;
    const Dev d = to_dev(q);
    if (tb) {
        int parity = 0;
        while (n >= n_lo) {
            const int ns = (n - n_lo + 1 < p->tb_k) ? n - n_lo + 1 : p->tb_k;
            rc = launch_tb(q, d, true, parity, n, 1, ns, st);
            if (rc) return rc;
            parity ^= 1;
            n -= ns;
        }
        if ((rc = tb_settle(q, parity, st))) return rc;
        if (defer) return FGPR_OK;
        const int64_t ncell = (int64_t)p->NI * p->NJ;
        hipLaunchKernelGGL(k_reduce_shots, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0, st,
                           p->gshot, p->S, ncell, p->grad_eps, p->grad_sig);
        HIP_TRY(hipGetLastError());
        return FGPR_OK;
    }
    for (; n >= n_lo; --n) {
        rc = launch_adjoint_step(d, q, n, 1, st);
        if (rc) return rc;
    }
    return FGPR_OK;
}

int fgpr_normalise(const double* g, int64_t n, double scale, float* out, void* stream) {
    if (!g || !out || n < 1) return fail(FGPR_EINVAL, "bad normalise arguments");
    hipLaunchKernelGGL(k_normalise, dim3(1), dim3(1024), 0, (hipStream_t)stream, g, n, scale, 1, out);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_profile_kernel(const fgpr_problem* p, int kernel, int n, int reps, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (n < 0 || n >= p->nt || reps < 1) return fail(FGPR_EINVAL, "bad step or repetition count");
    const bool graph = (kernel & FGPR_PROFILE_GRAPH) != 0;
    kernel &= ~FGPR_PROFILE_GRAPH;
    fgpr_problem q = *p;
    q.ctr = nullptr;
    const Dev d = to_dev(q);
    const dim3 grid = cell_grid(q), blk(kBlockJ, kBlockI, 1);
    hipStream_t st = (hipStream_t)stream;
    const int ninj = q.S * q.nrx;
    switch (kernel) {        // argument checks, before anything is launched or captured
        case FGPR_KERNEL_FWD_H:
            if (!q.traces && q.nrx > 0) return fail(FGPR_EINVAL, "null traces");
            break;
        case FGPR_KERNEL_FWD_E:
            if (q.nsrc > 0 && (!q.src || !q.srcscale)) return fail(FGPR_EINVAL, "null source arrays");
            break;
        case FGPR_KERNEL_ADJ_INJECT:
        case FGPR_KERNEL_ADJ_H:
            break;
        case FGPR_KERNEL_ADJ_E:
            if (n < 1 || !q.snaps || !q.grad_eps || !q.grad_sig)
                return fail(FGPR_EINVAL, "adjoint E kernel needs n >= 1, snaps and gradients");
            break;
        case FGPR_KERNEL_TB_FWD:
        case FGPR_KERNEL_TB_ADJ: {
            const bool adj = kernel == FGPR_KERNEL_TB_ADJ;
            if ((rc = tb_check(q, adj))) return rc;
            if (adj && (n < q.tb_k || !q.snaps)) return fail(FGPR_EINVAL, "adjoint block needs n >= tb_k and snaps");
            if (!adj && n + q.tb_k > q.nt) return fail(FGPR_EINVAL, "forward block runs past nt");
            break;
        }
        default:
            return fail(FGPR_EINVAL, "unknown kernel id");
    }
    auto body = [&](hipStream_t s) {
        for (int r = 0; r < reps; ++r) {
            switch (kernel) {
                case FGPR_KERNEL_FWD_H:
                    hipLaunchKernelGGL(k_h_phase<true>, grid, blk, 0, s, d, n, 1);
                    break;
                case FGPR_KERNEL_FWD_E:
                    if (q.snaps)
                        hipLaunchKernelGGL((k_e_phase<true, true, false>), grid, blk, 0, s, d, n, 1);
                    else
                        hipLaunchKernelGGL((k_e_phase<true, false, false>), grid, blk, 0, s, d, n, 1);
                    break;
                case FGPR_KERNEL_ADJ_INJECT:
                    if (ninj > 0 && q.resid)
                        hipLaunchKernelGGL(k_adj_inject, dim3((ninj + 255) / 256), dim3(256), 0, s, d, n, 1);
                    break;
                case FGPR_KERNEL_ADJ_E:
                    hipLaunchKernelGGL((k_e_phase<false, false, true>), grid, blk, 0, s, d, n, 1);
                    break;
                case FGPR_KERNEL_ADJ_H:
                    hipLaunchKernelGGL(k_h_phase<false>, grid, blk, 0, s, d, n, 1);
                    break;
                default: {       // the time-blocked blocks, checked above
                    const int rc2 = launch_tb(q, d, kernel == FGPR_KERNEL_TB_ADJ, r & 1, n, 1, q.tb_k, s);
                    if (rc2) return rc2;
                }
            }
        }
        HIP_TRY(hipGetLastError());
        return FGPR_OK;
    };
    if (!graph) return body(st);
    GraphKey key = make_key(q, 100 + kernel, reps);
    key.v.push_back(n);
    return launch_cached(g_prof_graphs, kProfGraphLimit, std::move(key), body, 1, st);
}

int fgpr_scale(const double* g, int64_t n, double scale, float* out, void* stream) {
    if (!g || !out || n < 1) return fail(FGPR_EINVAL, "bad scale arguments");
    hipLaunchKernelGGL(k_normalise, dim3(1), dim3(1024), 0, (hipStream_t)stream, g, n, scale, 0, out);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int64_t fgpr_tb_table_bytes(int nitems) {
    return nitems < 1 ? 0 : (int64_t)(tb_table_header(nitems) + (size_t)nitems * sizeof(MultiEntry));
}

int fgpr_tb_pack(const fgpr_tb_item* items, int nitems, int adj, void* host_table, int64_t bytes,
                 int* total_wgs) {
    if (!items || nitems < 1 || !host_table || !total_wgs) return fail(FGPR_EINVAL, "bad pack arguments");
    if (bytes < fgpr_tb_table_bytes(nitems)) return fail(FGPR_EINVAL, "table buffer too small");
    const fgpr_problem* p0 = items[0].prob;
    if (!p0) return fail(FGPR_EINVAL, "null problem in item 0");
    if (!tb_multi_kernel(p0->tb_ri, p0->tb_nt, p0->tb_k, adj != 0))
        return fail(FGPR_EINVAL, "variant not compiled for multi-problem launches");
    unsigned char* t = (unsigned char*)host_table;
    int* prefix = (int*)t;
    MultiEntry* ents = (MultiEntry*)(t + tb_table_header(nitems));
    int64_t acc = 0;
    for (int e = 0; e < nitems; ++e) {
        const fgpr_problem* p = items[e].prob;
        int rc = check_problem(p);
        if (rc) return rc;
        if (p->tb_ri != p0->tb_ri || p->tb_nt != p0->tb_nt || p->tb_k != p0->tb_k)
            return fail(FGPR_EINVAL, "all items of one launch must share the time-blocking variant");
        if ((rc = tb_check(*p, adj != 0))) return rc;
        const int n0 = items[e].n0, ns = items[e].nsteps;
        if (ns < 1 || ns > p->tb_k) return fail(FGPR_EINVAL, "nsteps must be in [1, tb_k]");
        if (!adj && (n0 < 0 || n0 + ns > p->nt)) return fail(FGPR_EINVAL, "forward block outside [0, nt)");
        if (adj && (n0 >= p->nt || n0 - ns + 1 < 1)) return fail(FGPR_EINVAL, "adjoint block outside [1, nt)");
        if (adj && !p->snaps) return fail(FGPR_EINVAL, "adjoint block needs snapshots");
        fgpr_problem q = *p;
        q.ctr = nullptr;
        prefix[e] = (int)acc;
        acc += (int64_t)p->tb_ti * p->tb_tj * p->S;
        if (acc > 0x7fffffff) return fail(FGPR_EINVAL, "too many workgroups in one launch");
        MultiEntry me;
        std::memset(&me, 0, sizeof(me));
        me.a = make_tb(q, to_dev(q), items[e].parity & 1, adj != 0);
        me.n0 = n0;
        me.nsteps = ns;
        std::memcpy(&ents[e], &me, sizeof(me));
    }
    prefix[nitems] = (int)acc;
    *total_wgs = (int)acc;
    return FGPR_OK;
}

int fgpr_tb_launch_table(const void* dev_table, int nitems, int total_wgs, int tb_ri, int tb_nt,
                         int tb_k, int adj, void* stream) {
    TBMultiKernel kfn = tb_multi_kernel(tb_ri, tb_nt, tb_k, adj != 0);
    if (!kfn) return fail(FGPR_EINVAL, "variant not compiled for multi-problem launches");
    if (!dev_table || nitems < 1 || total_wgs < 1) return fail(FGPR_EINVAL, "bad launch arguments");
    hipLaunchKernelGGL(kfn, dim3(total_wgs), dim3(tb_nt), 0, (hipStream_t)stream,
                       (const unsigned char*)dev_table, nitems);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_tb_settle(const fgpr_problem* p, int parity, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!p->Ez2 || !p->Hx2 || !p->Hy2 || (p->P > 0 && !p->psi2)) return fail(FGPR_EINVAL, "no partner buffers");
    return tb_settle(*p, parity & 1, (hipStream_t)stream);
}

int fgpr_reduce_shots(const fgpr_problem* p, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!p->gshot || !p->grad_eps || !p->grad_sig) return fail(FGPR_EINVAL, "reduce needs gshot and gradients");
    const int64_t ncell = (int64_t)p->NI * p->NJ;
    hipLaunchKernelGGL(k_reduce_shots, dim3((unsigned)((ncell + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, p->gshot, p->S, ncell, p->grad_eps, p->grad_sig);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_row_sums(const float* a, const float* b, int n, int stride, float* out, void* stream) {
    if (!a || !b || !out || n < 1 || stride < 1) return fail(FGPR_EINVAL, "bad row-sum arguments");
    hipLaunchKernelGGL(k_row_sums, dim3(1), dim3(64), 0, (hipStream_t)stream, a, b, n, stride, out);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int64_t fgpr_tv_workspace_bytes(int M, int N) {
    if (M < 1 || N < 1) return -1;
    return tv_bytes(M, N);
}

// The TV launches of one call.  `cache`: replay them from a captured graph
// (kept in a cache of its own, keyed by the pointers, sizes and parameters)
// unless the stream is being captured by the caller; otherwise launch them
// one by one (the allocation-free entry point: instantiating a graph
// allocates).
static int tv_run(const float* data, int M, int N, float lamda, float rho, int num_iter, float tol,
                  void* workspace, int64_t bytes, float* out, hipStream_t st, bool cache) {
    if (!data || !out || !workspace) return fail(FGPR_EINVAL, "null pointer");
    if (M < 1 || N < 1 || num_iter < 1 || !(rho > 0.f)) return fail(FGPR_EINVAL, "bad TV arguments");
    if (bytes < tv_bytes(M, N)) return fail(FGPR_EINVAL, "TV workspace too small");
    if ((uintptr_t)workspace % 256) return fail(FGPR_EINVAL, "TV workspace not 256-byte aligned");
    TVWork w;
    w.M2 = M + 2; w.N2 = N + 2;
    const int64_t plane = (int64_t)w.M2 * w.N2 * 4, pl = (plane + 255) / 256 * 256;
    char* base = (char*)workspace;
    float** planes[7] = {&w.X[0], &w.X[1], &w.Y, &w.Zx, &w.Zy, &w.Ux, &w.Uy};
    for (int k = 0; k < 7; ++k) *planes[k] = (float*)(base + k * pl);
    w.scal = (float*)(base + 7 * pl);
    w.flags = (int*)(base + 7 * pl + 64);
    w.part = (double*)(base + 7 * pl + 256);
    const dim3 blk(kTVBx, kTVBy), grid((w.N2 + kTVBx - 1) / kTVBx, (w.M2 + kTVBy - 1) / kTVBy);
    // float32 constants as the reference forms them: lamda * rho and
    // 1 + 4 * lamda * rho in double, then rounded; 1 / rho likewise
    const float lr = (float)((double)lamda * (double)rho);
    const float den = (float)(1.0 + 4.0 * (double)lamda * (double)rho);
    const float inv_rho = (float)(1.0 / (double)rho);
    const int nparts = (int)tv_blocks(M, N);
    auto body = [&](hipStream_t s) {
        hipLaunchKernelGGL(k_tv_max, dim3(1), dim3(1024), 0, s, data, (int64_t)M * N, w);
        hipLaunchKernelGGL(k_tv_init, grid, blk, 0, s, data, w);
        for (int k = 0; k < num_iter; ++k) {
            hipLaunchKernelGGL(k_tv_x, grid, blk, 0, s, w, k, lamda, lr, den, rho);
            hipLaunchKernelGGL(k_tv_zu, grid, blk, 0, s, w, k, rho, inv_rho, nparts, (double)tol);
        }
        hipLaunchKernelGGL(k_tv_out, dim3((N + kTVBx - 1) / kTVBx, (M + kTVBy - 1) / kTVBy), blk, 0,
                           s, w, num_iter, out);
        HIP_TRY(hipGetLastError());
        return FGPR_OK;
    };
    if (cache) {
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(st, &cap));
        cache = cap == hipStreamCaptureStatusNone;
    }
    if (!cache) return body(st);
    GraphKey key;
    auto I = [&](int64_t x) { key.v.push_back(x); };
    auto F = [&](float x) { int32_t b; std::memcpy(&b, &x, 4); key.v.push_back(b); };
    I((int64_t)(uintptr_t)data); I((int64_t)(uintptr_t)workspace); I((int64_t)(uintptr_t)out);
    I(M); I(N); F(lamda); F(rho); I(num_iter); F(tol);
    return launch_cached(g_tv_graphs, kTVGraphLimit, std::move(key), body, 1, st);
}

int fgpr_tv(const float* data, int M, int N, float lamda, float rho, int num_iter, float tol,
            void* workspace, int64_t bytes, float* out, void* stream) {
    return tv_run(data, M, N, lamda, rho, num_iter, tol, workspace, bytes, out, (hipStream_t)stream,
                  true);
}

int64_t fgpr_tb_map(const fgpr_problem* p, int adj, int32_t* shot, int32_t* tile, int8_t* interior,
                    int64_t n) {
    int rc = check_problem(p);
    if (rc) return rc;
    if ((rc = tb_check(*p, adj != 0))) return rc;
    const int64_t nwg = (int64_t)p->tb_ti * p->tb_tj * p->S;
    if (!shot || !tile || !interior || n < nwg) return fail(FGPR_EINVAL, "output arrays shorter than the launch");
    const TBArgs a = make_tb(*p, to_dev(*p), 0, adj != 0);
    for (int64_t wg = 0; wg < nwg; ++wg) {
        int s_ = 0, t_ = 0;
        interior[wg] = tb_item(a, p->S, (int)wg, (int)nwg, s_, t_) ? 1 : 0;
        shot[wg] = s_;
        tile[wg] = t_;
    }
    return nwg;
}

int64_t fgpr_fastdiv_mismatches(int d, int64_t n0, int64_t n1) {
    if (d < 1 || n0 < 0 || n1 > ((int64_t)1 << 31) || n0 > n1) return -1;
    const FastDiv f = make_fastdiv(d);
    int64_t bad = 0;
    for (int64_t n = n0; n < n1; ++n)
        bad += fdiv((int)n, f) != (int)(n / d);
    return bad;
}

int fgpr_tv_static(const float* data, int M, int N, float lamda, float rho, int num_iter, float tol,
                   float* out, void* stream) {
    if (M < 1 || N < 1) return fail(FGPR_EINVAL, "bad TV arguments");
    const int64_t need = tv_bytes(M, N);
    if (need > (int64_t)FGPR_TV_STATIC_BYTES)
        return fail(FGPR_EINVAL, "fgpr_tv_static: a " + std::to_string(M) + " x " + std::to_string(N) +
                                     " model needs " + std::to_string(need) +
                                     " bytes of workspace, more than the static " +
                                     std::to_string((int64_t)FGPR_TV_STATIC_BYTES) +
                                     "; use fgpr_tv with a caller workspace");
    void* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    // every word of the workspace the kernels read is written first in the
    // same call (k_tv_max: the maximum and the flags; k_tv_init: the seven
    // planes; each X sweep: its block sums), so nothing is cleared here
    return tv_run(data, M, N, lamda, rho, num_iter, tol, ws, (int64_t)FGPR_TV_STATIC_BYTES, out,
                  (hipStream_t)stream, false);
}

int64_t fgpr_snap_slot_floats(const fgpr_problem* p) {
    if (!p) return -1;
    if (!p->snap_tiled) return (int64_t)p->S * p->NI * p->pitch;
    if (!tb_kernel(p->tb_ri, p->tb_nt, p->tb_k, false)) return -1;
    return (int64_t)p->S * p->tb_ti * p->tb_tj * tb_snap_block(p->tb_ri, p->tb_nt, p->tb_k);
}

int64_t fgpr_tb_snap_floats(int NI, int NJ, int tb_ri, int tb_nt, int tb_k) {
    if (!tb_kernel(tb_ri, tb_nt, tb_k, false) || NI < 1 || NJ < 1) return -1;
    const int TI = tb_ri - 2 * tb_k - 1, TJ = kTBCols - 2 * tb_k - 1;
    const int64_t tiles = (int64_t)((NI + 1 + TI - 1) / TI) * ((NJ + 1 + TJ - 1) / TJ);
    return tiles * tb_snap_block(tb_ri, tb_nt, tb_k);
}

int fgpr_snap_pack(const fgpr_problem* p, int slot, const float* Ez, void* stream) {
    int rc = check_problem(p);
    if (rc) return rc;
    if (!p->snap_tiled || !p->snaps || !Ez) return fail(FGPR_EINVAL, "snap_pack needs a tiled snapshot buffer and Ez");
    if ((rc = tb_check(*p, false))) return rc;
    if (slot < 0 || slot >= p->nsnap) return fail(FGPR_EINVAL, "snapshot slot outside the buffer");
    const int64_t n = (int64_t)p->S * p->tb_ti * p->tb_tj * tb_snap_block(p->tb_ri, p->tb_nt, p->tb_k);
    hipLaunchKernelGGL(k_snap_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       to_dev(*p), p->tb_ri, p->tb_nt, p->tb_k, p->tb_ti, p->tb_tj, slot, Ez);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_fir(const float* x, int S, int nt, int nrx, const float* taps, int L, float* out,
             void* stream) {
    if (!x || !taps || !out) return fail(FGPR_EINVAL, "null pointer");
    if (S < 1 || nt < 1 || nrx < 1 || L < 1) return fail(FGPR_EINVAL, "bad FIR shape");
    if (L - 1 >= nt) return fail(FGPR_EINVAL, "FIR reflect padding (taps - 1) must be shorter than the trace");
    if (x == out) return fail(FGPR_EINVAL, "FIR input and output must not alias");
    const int64_t n = (int64_t)S * nt * nrx;
    hipLaunchKernelGGL(k_fir, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       x, S, nt, nrx, taps, L, out);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_release_graphs(void) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto& kv : g_graphs) destroy_entry(kv.second);
    g_graphs.clear();
    for (auto& kv : g_prof_graphs) destroy_entry(kv.second);
    g_prof_graphs.clear();
    for (auto& kv : g_tv_graphs) destroy_entry(kv.second);
    g_tv_graphs.clear();
    return FGPR_OK;
}

int fgpr_graph_count(void) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    return (int)g_graphs.size();
}

int fgpr_set_graph_limit(int n) {
    if (n < 1) return fail(FGPR_EINVAL, "graph cache limit must be >= 1");
    std::lock_guard<std::mutex> lk(g_graph_mu);
    const int prev = (int)g_graph_limit;
    g_graph_limit = (size_t)n;
    evict_over_limit(g_graphs, g_graph_limit, nullptr);
    return prev;
}

}  // extern "C"
