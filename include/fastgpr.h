/*
 * fastgpr.h -- C ABI of libfastgpr.so, the MI355X (gfx950) 2-D TMz FDTD
 * forward / adjoint solver with CPML, batched over shots.
 *
 * Plain C: pointers, sizes and one descriptor struct.  Every pointer inside
 * ``fgpr_problem`` is a DEVICE pointer owned by the caller (the Python host
 * allocates them as torch tensors); the library allocates no device memory
 * except captured HIP graphs.  Every call is asynchronous on the HIP stream it
 * is given (``stream`` = hipStream_t, NULL = default stream), returns 0 on
 * success and a negative code on error, and leaves a message readable through
 * fgpr_last_error() on the calling thread.
 *
 * Index conventions follow the reference: ``NI`` = G.nx (depth, first axis,
 * index i), ``NJ`` = G.ny (lateral, second axis, index j).  Fields are padded
 * to (NI+1) x (NJ+1) like the reference (compute.py:499-518) but the dead
 * k = 1 plane is gone and rows are padded to ``pitch`` floats
 * (fgpr_pitch(NJ)) for aligned, coalesced rows.
 *
 * The reference binds six CUDA libraries through ctypes (src/__init__.py:30-221)
 * and calls 12 launchers per time step.  This ABI replaces them with:
 *
 *   fgpr_cell_coeffs      <- Ucget  (uc.cu:48-55)  and Ucgeta (uc.cu:94-101)
 *   fgpr_forward          <- per step: launch_store_outputs (sourcereceiver.cu:40),
 *                            h_fields_updates (fields_updates_gpu.cu:80),
 *                            order1_{xminus,yminus,xplus,yplus} of pml_updates_h.cu
 *                            (:84,:314,:196,:431), e_fields_updates
 *                            (fields_updates_gpu.cu:35), order1_* of
 *                            pml_updates_e.cu (:80,:290,:173,:405),
 *                            update_hertzian_dipole (sourcereceiver.cu:89) and the
 *                            snapshot copy (compute.py:708-712); the time loop
 *                            of compute.py:574-712 runs inside the call.
 *   fgpr_adjoint          <- per step: back_source (back.cu:39), e_fields_updates,
 *                            order1_* (E), h_fields_updates, order1_* (H) and the
 *                            gradient accumulation of compute.py:1167-1172;
 *                            the reverse loop of compute.py:1048-1172.
 *   fgpr_normalise        <- grad / grad.norm() (compute.py:1174-1180).
 */
#ifndef FASTGPR_H
#define FASTGPR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGPR_ABI_VERSION 13

/* Error codes. */
#define FGPR_OK 0
#define FGPR_EINVAL (-1)   /* bad argument / inconsistent descriptor */
#define FGPR_EHIP (-2)     /* HIP runtime error */

/* Execution modes for fgpr_forward / fgpr_adjoint. */
#define FGPR_EXEC_EAGER 0  /* one kernel sequence per time step, launched from the host loop */
#define FGPR_EXEC_GRAPH 1  /* time loop replayed from captured HIP graphs of `chunk` steps */

/* Gradient modes for fgpr_adjoint. */
#define FGPR_GRAD_REFERENCE 0  /* reference's approximate gradient (compute.py:1167-1172) */
#define FGPR_GRAD_EXACT 1      /* exact discrete adjoint of the forward map (no reference counterpart) */

/*
 * One batch of S shots sharing one model and one grid.
 *
 * Layouts (all row-major, `pitch` = fgpr_pitch(NJ) floats per row):
 *   coef      5 planes of (NI+1)*pitch floats: uE0, uE1, uE4, uH1, uH4.
 *             Zero on i = NI, j = NJ and the pitch padding (uc.cu:19).
 *             uH0 is identically 1 inside and 0 on the padding, so it is
 *             not stored.
 *   pml       8*P floats: ERA, ERB, ERE, ERF, HRA, HRB, HRE, HRF (compute.py:211-222).
 *   src       S*nsrc*2 int32 cell indices (i, j); i < 0 marks a dead source.
 *   srcscale  nsrc*nt floats: w[n]*dl/(dx*dy*dz) (sourcereceiver.cu:73-75).
 *   rx        S*nrx*2 int32 cell indices (i, j); i < 0 marks a dead receiver.
 *   Ez,Hx,Hy  S*(NI+1)*pitch floats each.
 *   psi       fgpr_psi_floats(NI, NJ, P, S) floats, 8 slabs in this order:
 *             E-x0, E-xmax, H-x0, H-xmax  (each S*P*NJ, element [s][q][j]),
 *             E-y0, E-ymax, H-y0, H-ymax  (each S*NI*P, element [s][i][q]).
 *   traces    S*nt*nrx floats, element [s][n][r]        (forward output)
 *   snaps     nsnap*S*NI*pitch floats, element [slot(n)][s][i][j] with
 *             slot(n) = n - snap_base, or n % snap_mod when snap_mod > 0
 *             (forward output, adjoint input); NULL = do not store.
 *   resid     S*nt*nrx floats: dL/dtraces (adjoint input, back.cu:30).
 *   grad_eps, grad_sig   NI*NJ doubles, accumulated (+=) by fgpr_adjoint.  The
 *                        time-blocked reference adjoint accepts both null: it
 *                        then leaves its shots' sums in gshot for a later
 *                        fgpr_reduce_shots (shot groups of one batch on
 *                        separate streams, reduced once in shot order).
 *   ctr       one device int32 used as the step counter of captured graphs.
 */
typedef struct fgpr_problem {
    int32_t NI, NJ, P, S, nt, nsrc, nrx, pitch;
    float dx, dt;
    const float* coef;
    const float* pml;
    const int32_t* src;
    const float* srcscale;
    const int32_t* rx;
    float* Ez;
    float* Hx;
    float* Hy;
    float* psi;
    float* traces;
    float* snaps;
    const float* resid;
    double* grad_eps;
    double* grad_sig;
    int32_t* ctr;
    /* Time-blocked path, used when tb_k > 0 (see fgpr_tb_tiles). */
    float* Ez2;                 /* ping-pong partners of Ez, Hx, Hy, psi (same layouts) */
    float* Hx2;
    float* Hy2;
    float* psi2;
    int32_t tb_ri, tb_nt, tb_k; /* region rows, threads per workgroup, steps per launch */
    int32_t tb_ti, tb_tj;       /* tiles along i and j */
    int32_t tb_list_cap;        /* entries per (shot, tile) receiver block: 64 or 128 */
    const int32_t* tb_rx_list;  /* forward: per (shot, tile) a block of tb_list_cap entries, the
                                   receivers the tile centre samples in receiver order, then
                                   0xFFFFFFFF; entry = li << 26 | lj << 20 | r with (li, lj) the
                                   receiver's cell relative to the region's first row and column
                                   (tile * (TI, TJ) - tb_k) and r its index in the shot */
    const int32_t* tb_inj_list; /* adjoint: the same blocks for the receivers inside each region
                                   (centre and halo) */
    double* gshot;              /* adjoint: S x 2 x NI x NJ per-shot FP64 partial gradients, shot-major
                                   (a shot range is a contiguous slice); zeroed by the caller */
    float* scratch;             /* exact adjoint: 4*S*(NI+1)*pitch floats (no initialisation) */
    int32_t snap_base;          /* step of snapshot slot 0 (checkpointed segments; 0 otherwise) */
    int32_t nsnap;              /* snapshot slots in `snaps` (nt when it holds every step) */
    int32_t snap_mod;           /* > 0: step n lives in slot n % snap_mod (a ring of segment
                                   slots; the same captured graph then serves every segment) */
    int32_t snap_tiled;         /* 1: snapshots in the time-blocked layout of this problem's
                                   variant -- per (slot, shot, tile) a block of the waves owning
                                   tile-centre rows x the tile-centre lanes x the rows of a
                                   thread, ordered [wave][lane][row] -- written by the time-blocked
                                   forward and read by the time-blocked adjoint of the same
                                   variant; 0: the row-major layout above */
} fgpr_problem;

/* ABI version (FGPR_ABI_VERSION). */
int fgpr_abi_version(void);

/* sizeof(fgpr_problem), for binding layers to check their mirror of it. */
size_t fgpr_problem_size(void);

/* Message of the last failed call on this thread ("" if none). */
const char* fgpr_last_error(void);

/* Row pitch, in floats, of every field / coefficient plane for NJ columns. */
int fgpr_pitch(int NJ);

/* Number of floats of the psi workspace for one batch. */
int64_t fgpr_psi_floats(int NI, int NJ, int P, int S);

/*
 * Tile geometry of a time-blocked variant: centre tile size (*TI x *TJ) and
 * tile counts for an NI x NJ grid.  Returns FGPR_EINVAL for a variant that
 * is not compiled in.  Supported (tb_ri, tb_nt): (32, 256), (64, 512),
 * (64, 256), (32, 512), (64, 1024); tb_k up to 8.
 */
int fgpr_tb_tiles(int NI, int NJ, int tb_ri, int tb_nt, int tb_k, int* TI, int* TJ,
                  int* tiles_i, int* tiles_j);

/* Check a descriptor without launching anything (0 = consistent). */
int fgpr_validate(const fgpr_problem* p);

/*
 * Per-cell update coefficients from compact (NI, NJ) float er, se, mr into
 * the 5-plane `coef` layout above.  adjoint = 0 reproduces Ucget
 * (uc.cu:11-55), adjoint = 1 reproduces Ucgeta (uc.cu:60-101).
 */
int fgpr_cell_coeffs(const float* er, const float* se, const float* mr,
                     int NI, int NJ, float dt, float dx, int adjoint,
                     float* coef, void* stream);

/*
 * Forward steps n = n0 .. n1-1 (compute.py:574-712).  Fields and psi must be
 * zero before step 0 (the caller clears them).  mode = FGPR_EXEC_EAGER or
 * FGPR_EXEC_GRAPH; chunk = steps per captured graph (ignored when eager).
 */
int fgpr_forward(const fgpr_problem* p, int n0, int n1, int mode, int chunk, void* stream);

/*
 * Adjoint steps n = n_hi down to n_lo (inclusive).
 * FGPR_GRAD_REFERENCE (n_lo >= 1): the reference's order (compute.py:1048-1172):
 *   inject -resid at the receivers, E + E-CPML, H + H-CPML with the Ucgeta
 *   coefficients, accumulation of sum (Ez[n]-Ez[n-1])*lambda into grad_eps and
 *   sum Ez[n]*lambda into grad_sig (FP64, summed over shots; the constant dt of
 *   the sigma term is left out -- it cancels in the normalisation).
 * FGPR_GRAD_EXACT (n_lo >= 0, pass the FORWARD coefficients): the transpose of
 *   the forward stepping; grad_eps accumulates sum uE4*lambda*(E^{n+1}-E^n)
 *   (dL/d er = -(e0/dt) times it) and grad_sig accumulates dL/d se directly.
 *   Needs `scratch`.  psi / fields hold the adjoint state and must start at 0.
 */
int fgpr_adjoint(const fgpr_problem* p, int n_hi, int n_lo, int grad_mode, int mode,
                 int chunk, void* stream);

/*
 * out[k] = (float)(scale * g[k] / ||g||_2) over n doubles (compute.py:1174-1180).
 * Uses one deterministic single-workgroup reduction.
 */
int fgpr_normalise(const double* g, int64_t n, double scale, float* out, void* stream);

/* Kernel ids for fgpr_profile_kernel. */
#define FGPR_KERNEL_FWD_H 0      /* H + H-CPML + receiver sampling */
#define FGPR_KERNEL_FWD_E 1      /* E + E-CPML + source + snapshot store */
#define FGPR_KERNEL_ADJ_INJECT 2 /* adjoint source at the receivers */
#define FGPR_KERNEL_ADJ_E 3      /* E + E-CPML + gradient accumulation */
#define FGPR_KERNEL_ADJ_H 4      /* H + H-CPML (adjoint coefficients) */
#define FGPR_KERNEL_TB_FWD 5     /* time-blocked forward block of tb_k steps from step n */
#define FGPR_KERNEL_TB_ADJ 6     /* time-blocked adjoint block of tb_k steps down from step n */

/* Or'd into fgpr_profile_kernel's `kernel`: replay the launches from a graph. */
#define FGPR_PROFILE_GRAPH 0x100

/*
 * Launch one solver kernel `reps` times back to back at step n, on the
 * descriptor's buffers, for timing with stream events.  It advances the
 * state like the real step does, so run it on scratch buffers.  With
 * FGPR_PROFILE_GRAPH the `reps` launches are captured once into a HIP graph
 * (kept in the sweep-graph cache, keyed by the descriptor, kernel, step and
 * reps) and each call replays it, as the sweeps replay theirs: the first
 * call with a key also captures, so time the second.  No counterpart in the
 * reference (measurement only).
 */
int fgpr_profile_kernel(const fgpr_problem* p, int kernel, int n, int reps, void* stream);

/* out[k] = (float)(scale * g[k]) over n doubles. */
int fgpr_scale(const double* g, int64_t n, double scale, float* out, void* stream);

/*
 * Multi-problem time-blocked launches (uneven batches, serving).  One launch
 * advances many problems -- different grids, shot counts and time steps --
 * by up to tb_k steps each; all items share (tb_ri, tb_nt, tb_k).  No
 * counterpart in the reference (it runs one model per call).
 *   fgpr_tb_pack        builds the launch table in host memory (validates
 *                       every item; *total_wgs = workgroups of the launch);
 *   fgpr_tb_launch_table launches it from a device copy of that table;
 *   fgpr_tb_settle       copies a problem's state back from the partner
 *                       buffers after an odd number of its blocks;
 *   fgpr_reduce_shots    grad_eps/grad_sig += sum over shots of gshot, in shot
 *                       order (also the deferred reduction of shot groups,
 *                       see grad_eps above).
 * Supported variants: (32,512,{2,4,6}), (64,1024,{2,4,6,8}), (64,512,{2,4,6,8}).
 */
typedef struct fgpr_tb_item {
    const fgpr_problem* prob;   /* host pointer to the problem descriptor */
    int32_t n0;                 /* first step (forward: ascending, adjoint: descending) */
    int32_t nsteps;             /* steps of this block, 1..tb_k */
    int32_t parity;             /* 0: state in Ez/Hx/Hy/psi, 1: in the partner buffers */
    int32_t pad;
} fgpr_tb_item;

int64_t fgpr_tb_table_bytes(int nitems);
int fgpr_tb_pack(const fgpr_tb_item* items, int nitems, int adj, void* host_table, int64_t bytes,
                 int* total_wgs);
int fgpr_tb_launch_table(const void* dev_table, int nitems, int total_wgs, int tb_ri, int tb_nt,
                         int tb_k, int adj, void* stream);
int fgpr_tb_settle(const fgpr_problem* p, int parity, void* stream);
int fgpr_reduce_shots(const fgpr_problem* p, void* stream);

/* Total-variation denoiser, replaces total_variation (src/tv.py:135-217) as
 * called from GPRFWI.backward (compute.py:916-933): ADMM on data / (max + 1e-6)
 * zero-padded by one cell, lamda, rho, at most num_iter iterations, stopping
 * when ||X - X_prev||_2 < tol; out = X * max.  data and out are M*N floats
 * (row-major, the model without its trailing unit axis).  The iterations run
 * on the device without host synchronisation; workspace is caller-owned
 * device memory of fgpr_tv_workspace_bytes(M, N) bytes, 256-byte aligned.
 * The call's launches (about two per iteration) are captured once into a
 * HIP graph, cached by (data, workspace, out, M, N, lamda, rho, num_iter,
 * tol) in a cache of at most 8, and replayed; a caller that keeps those
 * buffers in place replays one graph per call.  On a stream the caller is
 * capturing, the launches go into the caller's capture instead. */
int64_t fgpr_tv_workspace_bytes(int M, int N);
int fgpr_tv(const float* data, int M, int N, float lamda, float rho, int num_iter, float tol,
            void* workspace, int64_t bytes, float* out, void* stream);

/* fgpr_tv without a workspace argument: the workspace is a static device
 * array of FGPR_TV_STATIC_BYTES inside the library (one per device, held
 * until the process ends), so the call allocates nothing -- for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * Models whose fgpr_tv_workspace_bytes(M, N) exceeds the static size (about
 * 2.39 million padded cells, e.g. 2000 x 1000) are rejected with
 * FGPR_EINVAL.  Calls on one device share the array: calls that overlap in
 * time on one device (two streams, two threads) race and must be ordered by
 * the caller.  Same results as fgpr_tv, bitwise.  It launches its kernels
 * one by one (no graph is captured or instantiated). */
#define FGPR_TV_STATIC_BYTES (64u << 20)
int fgpr_tv_static(const float* data, int M, int N, float lamda, float rho, int num_iter, float tol,
                   float* out, void* stream);

/* out[0] = a[0] + a[s] + a[2s] + ..., out[1] the same over b (n terms,
 * stride s, FP32, in index order): the sequential sums of the first model
 * row that set the CPML sigma_max (compute.py:278-281), on the device so
 * that a call needs no device-to-host copy of the model. */
int fgpr_row_sums(const float* a, const float* b, int n, int stride, float* out, void* stream);

/* Floats per snapshot slot in the problem's layout (row-major S*NI*pitch, or
 * S * fgpr_tb_snap_floats(...) when snap_tiled). */
int64_t fgpr_snap_slot_floats(const fgpr_problem* p);

/* Floats per shot of one snapshot slot in the time-blocked layout of variant
 * (tb_ri, tb_nt, tb_k) on an NI x NJ grid: per tile, the waves that own
 * tile-centre rows times the tile-centre lanes times the rows per thread
 * (about 1.1x NI*NJ for the 64-row variants); -1 for a variant that is not
 * compiled in.  No counterpart in the reference. */
int64_t fgpr_tb_snap_floats(int NI, int NJ, int tb_ri, int tb_nt, int tb_k);

/* The work-item mapping of one time-blocked launch of p (adj = 0 forward,
 * 1 adjoint), evaluated on the host by the code the kernels run: for every
 * workgroup wg < S * tiles, its shot, its tile and whether it is one of the
 * interior items the launch orders first.  Returns the workgroup count or a
 * negative code.  No GPU work; no counterpart in the reference (inspection
 * and tests). */
int64_t fgpr_tb_map(const fgpr_problem* p, int adj, int32_t* shot, int32_t* tile, int8_t* interior,
                    int64_t n);

/* Count of dividends n in [n0, n1) (0 <= n0 <= n1 <= 2^31) for which the
 * kernels' multiply-and-shift division by d (FastDiv, evaluated on the host)
 * differs from n / d; -1 for bad arguments.  No counterpart in the
 * reference (tests). */
int64_t fgpr_fastdiv_mismatches(int d, int64_t n0, int64_t n1);

/* Writes slot `slot` of a tiled snapshot buffer from a row-major Ez
 * (S*(NI+1)*pitch floats, the layout of p->Ez): seeds the ring slot before a
 * checkpointed segment is recomputed.  No counterpart in the reference. */
int fgpr_snap_pack(const fgpr_problem* p, int slot, const float* Ez, void* stream);

/* FIR low-pass of traces, replaces the conv1d of apply_filter
 * (multiscale.py:19-50): x and out are S*nt*nrx floats laid out [s][t][r]
 * (the layout of `traces`), taps has L floats (design_fir_filter,
 * multiscale.py:3-17); each trace is reflect-padded by L-1 samples on the left,
 * so out[s][t][r] = sum_k taps[k] * x[s][|t + k - (L-1)|][r].  Requires L-1 < nt
 * (the reference's padding raises otherwise) and out != x. */
int fgpr_fir(const float* x, int S, int nt, int nrx, const float* taps, int L, float* out,
             void* stream);

/* Release every HIP graph the library captured; call it only when none of
 * them is still running (after synchronising the device).  Graphs are cached
 * by the descriptor's fields (device pointers and sizes), direction and
 * chunk; the cache holds at most fgpr_set_graph_limit() executables (64 by
 * default) and destroys the least recently launched one beyond that, after
 * its last launch has finished, so callers whose buffers move between calls
 * keep a bounded number of graphs.  The graphs of fgpr_profile_kernel
 * (FGPR_PROFILE_GRAPH) and of fgpr_tv sit in caches of their own, of at most
 * 8 each, and are released here too. */
int fgpr_release_graphs(void);

/* Number of captured sweep graphs the cache holds now (profile graphs not
 * counted). */
int fgpr_graph_count(void);

/* Set the cache bound (n >= 1, evicting at once if over it); returns the
 * previous bound. */
int fgpr_set_graph_limit(int n);

#ifdef __cplusplus
}
#endif

#endif /* FASTGPR_H */
