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
// snapshot store are fused into the two half-step kernels.  Floating-point
// contraction is off in this file so each expression rounds exactly as the
// reference's FP32 operation order does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
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
    float dx;
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
    d.coef = p.coef; d.pml = p.pml; d.src = p.src; d.srcscale = p.srcscale; d.rx = p.rx;
    d.Ez = p.Ez; d.Hx = p.Hx; d.Hy = p.Hy; d.psi = p.psi;
    d.traces = p.traces; d.snaps = p.snaps; d.resid = p.resid;
    d.grad_eps = p.grad_eps; d.grad_sig = p.grad_sig; d.ctr = p.ctr;
    return d;
}

__device__ __forceinline__ size_t plane(const Dev& d) { return (size_t)(d.NI + 1) * d.pitch; }

__device__ __forceinline__ size_t xslab_floats(const Dev& d) { return (size_t)d.S * d.P * d.NJ; }
__device__ __forceinline__ size_t yslab_floats(const Dev& d) { return (size_t)d.S * d.NI * d.P; }

__device__ __forceinline__ float* psi_x(const Dev& d, int kind, int s, int q, int j) {
    return d.psi + (size_t)kind * xslab_floats(d) + ((size_t)s * d.P + q) * d.NJ + j;
}
__device__ __forceinline__ float* psi_y(const Dev& d, int kind, int s, int i, int q) {
    return d.psi + 4 * xslab_floats(d) + (size_t)(kind - 4) * yslab_floats(d) +
           ((size_t)s * d.NI + i) * d.P + q;
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
    const float dx = d.dx;

    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        const float ez = d.Ez[b];
        const float ezj = d.Ez[b + 1];       // (i, j+1), j+1 <= NJ
        const float ezi = d.Ez[b + pitch];   // (i+1, j), i+1 <= NI
        // Hx: main update for i >= 1 (uH0 == 1), y-CPML on every row.
        if (i >= 1 || yq >= 0) {
            float hx = d.Hx[b];
            if (i >= 1) hx = hx - uH1 * (ezj - ez);
            if (yq >= 0) {
                float* pp = psi_y(d, yk, s, i, yq);
                const float ps = *pp;
                const float dd = (ezj - ez) / dx;
                hx = hx - uH4 * (yRA01 * dd + yRB * ps);
                *pp = yRE * ps - yRF * dd;
            }
            d.Hx[b] = hx;
        }
        // Hy: main update for j >= 1, x-CPML on every column.
        if (j >= 1 || xq >= 0) {
            float hy = d.Hy[b];
            if (j >= 1) hy = hy + uH1 * (ezi - ez);
            if (xq >= 0) {
                float* pp = psi_x(d, xk, s, xq, j);
                const float ps = *pp;
                const float dd = (ezi - ez) / dx;
                hy = hy + uH4 * (xRA01 * dd + xRB * ps);
                *pp = xRE * ps - xRF * dd;
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
    const float dx = d.dx;
    const bool main_upd = (i >= 1 && j >= 1);
    double acc_e = 0.0, acc_s = 0.0;

    for (int s = 0; s < d.S; ++s) {
        const size_t b = (size_t)s * pl + o;
        float ez = d.Ez[b];
        const float hy = d.Hy[b];
        const float hx = d.Hx[b];
        const float hyim = (i >= 1) ? d.Hy[b - pitch] : 0.f;
        const float hxjm = (j >= 1) ? d.Hx[b - 1] : 0.f;
        if (main_upd) ez = (uE0 * ez + uE1 * (hy - hyim)) - uE1 * (hx - hxjm);
        // x-CPML corrects Ez through dHy/dx, y-CPML through dHx/dy.
        float xcorr_d = 0.f, ycorr_d = 0.f;
        if (xq >= 0) xcorr_d = (hy - hyim) / dx;
        if (yq >= 0) ycorr_d = (hx - hxjm) / dx;
        if (xk == EX0 && xq >= 0) {
            float* pp = psi_x(d, EX0, s, xq, j);
            const float ps = *pp;
            ez = ez + uE4 * (xRA01 * xcorr_d + xRB * ps);
            *pp = xRE * ps - xRF * xcorr_d;
        }
        if (yk == EY0 && yq >= 0) {
            float* pp = psi_y(d, EY0, s, i, yq);
            const float ps = *pp;
            ez = ez - uE4 * (yRA01 * ycorr_d + yRB * ps);
            *pp = yRE * ps - yRF * ycorr_d;
        }
        if (xk == EXM && xq >= 0) {
            float* pp = psi_x(d, EXM, s, xq, j);
            const float ps = *pp;
            ez = ez + uE4 * (xRA01 * xcorr_d + xRB * ps);
            *pp = xRE * ps - xRF * xcorr_d;
        }
        if (yk == EYM && yq >= 0) {
            float* pp = psi_y(d, EYM, s, i, yq);
            const float ps = *pp;
            ez = ez - uE4 * (yRA01 * ycorr_d + yRB * ps);
            *pp = yRE * ps - yRF * ycorr_d;
        }
        if (SRC) {
            for (int k = 0; k < d.nsrc; ++k) {
                const int* sc = d.src + 2 * ((size_t)s * d.nsrc + k);
                if (sc[0] == i && sc[1] == j) ez = ez - uE4 * d.srcscale[(size_t)k * d.nt + n];
            }
        }
        d.Ez[b] = ez;
        if (SNAP) d.snaps[(((size_t)n * d.S + s) * NI + i) * pitch + j] = ez;
        if (GRAD) {
            const size_t sn = (((size_t)n * d.S + s) * NI + i) * pitch + j;
            const size_t sp = sn - (size_t)d.S * NI * pitch;
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
__global__ void k_adj_inject(Dev d, int koff, int dir) {
    const int n = step_index(d, koff, dir);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.S * d.nrx) return;
    const int s = t / d.nrx, r = t - s * d.nrx;
    const int ri = d.rx[2 * t], rj = d.rx[2 * t + 1];
    if (ri < 0 || ri >= d.NI || rj >= d.NJ) return;
    const float g = d.resid[((size_t)s * d.nt + n) * d.nrx + r];
    atomicAdd(&d.Ez[((size_t)s * (d.NI + 1) + ri) * d.pitch + rj], -g);
}

__global__ void k_advance_ctr(int* ctr, int delta) {
    if (threadIdx.x == 0) ctr[0] = ctr[0] + delta;
}

__global__ void k_set_ctr(int* ctr, int v) {
    if (threadIdx.x == 0) ctr[0] = v;
}

// Deterministic single-workgroup sum of squares, then scaled FP32 output.
__global__ void __launch_bounds__(1024) k_normalise(const double* __restrict__ g, int64_t n,
                                                    double scale, float* __restrict__ out) {
    __shared__ double red[1024];
    double acc = 0.0;
    for (int64_t k = threadIdx.x; k < n; k += 1024) acc += g[k] * g[k];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    const double inv = scale / sqrt(red[0]);
    for (int64_t k = threadIdx.x; k < n; k += 1024) out[k] = (float)(g[k] * inv);
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

// Captured graphs: one per (descriptor contents, direction, chunk).  The
// descriptor holds the device pointers baked into the graph, so it is the
// cache key byte for byte.
struct GraphKey {
    std::vector<unsigned char> bytes;
    bool operator<(const GraphKey& o) const { return bytes < o.bytes; }
};

std::mutex g_graph_mu;
std::map<GraphKey, hipGraphExec_t> g_graphs;

GraphKey make_key(const fgpr_problem& p, int kind, int chunk) {
    GraphKey k;
    k.bytes.resize(sizeof(fgpr_problem) + 2 * sizeof(int));
    std::memcpy(k.bytes.data(), &p, sizeof(fgpr_problem));
    std::memcpy(k.bytes.data() + sizeof(fgpr_problem), &kind, sizeof(int));
    std::memcpy(k.bytes.data() + sizeof(fgpr_problem) + sizeof(int), &chunk, sizeof(int));
    return k;
}

// kind 0 = forward (counter advances by +chunk), kind 1 = adjoint (-chunk).
int get_graph(const fgpr_problem& p, int kind, int chunk, hipGraphExec_t* out) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    GraphKey key = make_key(p, kind, chunk);
    auto it = g_graphs.find(key);
    if (it != g_graphs.end()) {
        *out = it->second;
        return FGPR_OK;
    }
    hipStream_t cs;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    const Dev d = to_dev(p);
    const int dir = kind == 0 ? 1 : -1;
    hipGraph_t g;
    int rc = FGPR_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipStreamDestroy(cs);
        return fail(FGPR_EHIP, "hipStreamBeginCapture failed");
    }
    for (int k = 0; k < chunk && rc == FGPR_OK; ++k)
        rc = kind == 0 ? launch_forward_step(d, p, k, dir, cs) : launch_adjoint_step(d, p, k, dir, cs);
    hipLaunchKernelGGL(k_advance_ctr, dim3(1), dim3(64), 0, cs, p.ctr, dir * chunk);
    hipError_t ec = hipStreamEndCapture(cs, &g);
    (void)hipStreamDestroy(cs);
    if (rc != FGPR_OK) return rc;
    if (ec != hipSuccess) return fail(FGPR_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipGraphExec_t ex;
    ec = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ec != hipSuccess) return fail(FGPR_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
    g_graphs.emplace(std::move(key), ex);
    *out = ex;
    return FGPR_OK;
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
    return FGPR_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

int fgpr_abi_version(void) { return FGPR_ABI_VERSION; }

const char* fgpr_last_error(void) { return g_err.c_str(); }

int fgpr_pitch(int NJ) { return ((NJ + 1 + 63) / 64) * 64; }

int64_t fgpr_psi_floats(int NI, int NJ, int P, int S) {
    return 4 * (int64_t)S * P * NJ + 4 * (int64_t)S * NI * P;
}

int fgpr_validate(const fgpr_problem* p) { return check_problem(p); }

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
    hipStream_t st = (hipStream_t)stream;
    fgpr_problem q = *p;
    int n = n0;
    if (mode == FGPR_EXEC_GRAPH && chunk > 0 && p->ctr && n1 - n0 >= chunk) {
        hipGraphExec_t ex;
        rc = get_graph(q, 0, chunk, &ex);
        if (rc) return rc;
        hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(64), 0, st, q.ctr, n0);
        for (; n + chunk <= n1; n += chunk) HIP_TRY(hipGraphLaunch(ex, st));
    }
    q.ctr = nullptr;
    const Dev d = to_dev(q);
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
    if (grad_mode != FGPR_GRAD_REFERENCE) return fail(FGPR_EINVAL, "unknown gradient mode");
    if (n_lo < 1 || n_hi >= p->nt || n_hi < n_lo - 1) return fail(FGPR_EINVAL, "bad step range");
    if (!p->snaps || !p->grad_eps || !p->grad_sig) return fail(FGPR_EINVAL, "adjoint needs snaps and gradients");
    if (p->nrx > 0 && (!p->resid || !p->rx)) return fail(FGPR_EINVAL, "null residual/receivers");
    hipStream_t st = (hipStream_t)stream;
    fgpr_problem q = *p;
    int n = n_hi;
    if (mode == FGPR_EXEC_GRAPH && chunk > 0 && p->ctr && n_hi - n_lo + 1 >= chunk) {
        hipGraphExec_t ex;
        rc = get_graph(q, 1, chunk, &ex);
        if (rc) return rc;
        hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(64), 0, st, q.ctr, n_hi);
        for (; n - chunk + 1 >= n_lo; n -= chunk) HIP_TRY(hipGraphLaunch(ex, st));
    }
    q.ctr = nullptr;
    const Dev d = to_dev(q);
    for (; n >= n_lo; --n) {
        rc = launch_adjoint_step(d, q, n, 1, st);
        if (rc) return rc;
    }
    return FGPR_OK;
}

int fgpr_normalise(const double* g, int64_t n, double scale, float* out, void* stream) {
    if (!g || !out || n < 1) return fail(FGPR_EINVAL, "bad normalise arguments");
    hipLaunchKernelGGL(k_normalise, dim3(1), dim3(1024), 0, (hipStream_t)stream, g, n, scale, out);
    HIP_TRY(hipGetLastError());
    return FGPR_OK;
}

int fgpr_release_graphs(void) {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    for (auto& kv : g_graphs) (void)hipGraphExecDestroy(kv.second);
    g_graphs.clear();
    return FGPR_OK;
}

}  // extern "C"
