/*
 * bfpn.h — C ABI of the MI355X-native in-core eq15 BFPn solver.
 *
 * The reference solver (BFPnSolver + four operator classes) is a C++/CUDA
 * API with raw device pointers, host-synchronous calls and CUDA_CHECK ->
 * exit(1) error handling.  This header replaces that seam with plain
 * extern "C" entry points: POD parameter structs, opaque handles, status
 * codes (never exit()), stream-ordered calls without hidden host syncs, and
 * no global state shared between contexts.
 *
 * Every entry point cites the reference interface it replaces.  Reference
 * paths are relative to the root of xiaojin1423764/BFPn_GPU.
 *
 * Layout of every phase-space array (unchanged from the reference,
 * src/bfpm_solver.cu:2013):
 *     idx = (g * nyz + s) * nang + m,  s = iz*(Ny+1) + iy,  m = iv*Nom + iu
 * with nyz = (Ny+1)*(Nz+1) transverse nodes and nang = Nmu*Nom directions.
 *
 * State precision: every operator exists for FP32 state ("_f32") and FP64
 * state ("_f64").  The energy recurrence and all dose reductions always
 * accumulate in FP64 whatever the state type.
 */
#ifndef BFPN_H
#define BFPN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BFPN_API __attribute__((visibility("default")))

/* ---- status codes (replace CUDA_CHECK -> exit(1), src/utils.cuh:13-21) ---- */
typedef enum bfpn_status {
    BFPN_OK = 0,
    BFPN_E_INVALID = 1,     /* bad argument / inconsistent sizes            */
    BFPN_E_NOMEM = 2,       /* device memory preflight or allocation failed */
    BFPN_E_HIP = 3,         /* HIP runtime error                            */
    BFPN_E_IO = 4,          /* table file missing or short                  */
    BFPN_E_UNSUPPORTED = 5  /* option not supported by this library         */
} bfpn_status;

typedef enum bfpn_precision {
    BFPN_FP32 = 0,          /* FP32 state, FP64 energy recurrence and IDD   */
    BFPN_FP64 = 1           /* FP64 everywhere (oracle-grade instantiation) */
} bfpn_precision;

/* GridParams (src/bfpm_solver.cuh:21-27); deltas are derived by the library
 * exactly as GridParams::computeDeltas (src/bfpm_solver.cu:1116-1126). */
typedef struct bfpn_grid {
    int Nx, Ny, Nz, Ng, Nmu, Nom;
    double Lx, Ly, Lz, Lu, Lv, Lg;
} bfpn_grid;

/* PhysicsParams (src/bfpm_solver.cuh:30-62), in-core fields only.  The
 * streaming/lite knobs of the reference are CLI-level aliases here. */
typedef struct bfpn_phys {
    double sig_E;
    double a_1, a_2;
    double beam_energy;
    double density;
    double energy_density;
    char material_name[32];          /* "", "water", "bone", "air"          */
    int primary_only;
    int eq15_straggling;
    int energy_only;
    int no_transport;
    int no_angle;
    int no_spatial_clipping;
    int discrete_angle_delta;
    int normalize_initial_mass;
    int trapezoidal_yz;
    int dose_catastrophic_loss;
    int calibrate_cross_sections;
    int save_energy_moments;
    int legacy_energy;               /* --energy-model legacy: first-order
                                        upwind energy step (src/energy_solver.cu:88-106) */
    int lite_memory;                 /* --lite-memory: primary psi^1 only, one
                                        upwind energy step per depth step
                                        (src/bfpm_solver.cu:2191-2212, 3396-3416) */
} bfpn_phys;

/* Derived deltas (GridParams::computeDeltas, src/bfpm_solver.cu:1116-1126). */
typedef struct bfpn_deltas {
    double dt, dy, dz, dg, du, dv;
} bfpn_deltas;

typedef struct bfpn_ctx bfpn_ctx;         /* one beam, all state in HBM */
typedef struct bfpn_workspace bfpn_workspace; /* scratch for operator calls */

/* Last error message of the calling thread ("" if none). */
BFPN_API const char* bfpn_last_error(void);
BFPN_API const char* bfpn_version(void);

/* ---------------------------------------------------------------------------
 * Device memory helpers.  Callers that hold no other HIP runtime (ctypes,
 * the CLI) use these to stage operator inputs.  Replaces DeviceArray<T>
 * (src/utils.cuh:55-140).
 * ------------------------------------------------------------------------- */
BFPN_API bfpn_status bfpn_device_count(int* count);
BFPN_API bfpn_status bfpn_device_info(int device, char* name, size_t name_len,
                                      size_t* total_bytes, int* cu_count);
BFPN_API bfpn_status bfpn_malloc(void** dptr, size_t bytes);
BFPN_API bfpn_status bfpn_free(void* dptr);
BFPN_API bfpn_status bfpn_memcpy_h2d(void* dst, const void* src, size_t bytes);
BFPN_API bfpn_status bfpn_memcpy_d2h(void* dst, const void* src, size_t bytes);
BFPN_API bfpn_status bfpn_memset(void* dptr, int value, size_t bytes);
/* Device-to-device copy of `bytes` (a multiple of 16) by a plain streaming
 * kernel, one read and one write of each byte; synchronous.  No counterpart
 * in the reference: it is the calibration kernel for HBM traffic counters. */
BFPN_API bfpn_status bfpn_device_copy(void* dst, const void* src, size_t bytes);
BFPN_API bfpn_status bfpn_synchronize(void);
BFPN_API bfpn_status bfpn_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---------------------------------------------------------------------------
 * Operator-level API (the seam of the reference's sub-solver classes).
 * All pointers are device pointers unless stated; `stream` is a hipStream_t
 * (NULL = the default stream).  Calls are stream-ordered: each call builds
 * its per-group coefficient tables on the device from the caller's device
 * tables and enqueues its kernels, without waiting on the host.  The host
 * waits (hipStreamSynchronize on `stream`) only when a call
 *   - returns a host value (bfpn_idd_*),
 *   - grows a workspace buffer (the first call at a larger size),
 *   - builds the source operands: bfpn_secondary_source_prepare_*, or a
 *     source call whose table pointers or sizes differ from the prepared ones,
 *   - is an angular call whose sig_tr values or grid scalars differ from the
 *     previous angular call's (the CN factors are built on the host).
 * Scratch comes from a workspace that grows on demand, like the sub-solvers'
 * lazily grown buffers (src/energy_solver.cu:461-464,
 * src/spatial_transport.cu:211-215, src/angle_diffusion.cu:429-440).  A
 * workspace orders its scratch on the stream of each call: calls on
 * different streams that share a workspace must be ordered by the caller.
 * ------------------------------------------------------------------------- */
BFPN_API bfpn_status bfpn_workspace_create(bfpn_workspace** ws);
BFPN_API void        bfpn_workspace_destroy(bfpn_workspace* ws);
/* Host-side builds of cached operands the workspace has made so far (source
 * operands, CN factors) and the device bytes it holds.  No counterpart in the
 * reference; it makes the caching contract above observable. */
BFPN_API bfpn_status bfpn_workspace_stats(const bfpn_workspace* ws, int* builds,
                                          size_t* device_bytes);

/* Eq.15 DG/CN energy half step, in place on (F, f).
 * Replaces EnergyDepositionSolver::solve with the eq15 model
 * (src/energy_solver.cuh:65-77, src/energy_solver.cu:41-161, 444-506).
 * S has Ng+1 entries, sigma and T have Ng.  Q1/Q2 may be NULL (primary);
 * T may be NULL or straggling = 0 to disable straggling. */
BFPN_API bfpn_status bfpn_energy_half_f32(
    bfpn_workspace* ws, float* F, float* f,
    const double* S, const double* sigma, const double* T,
    double h, double dg, double density, int Ng, int nyz, int nang,
    const float* Q1, const float* Q2, int straggling, void* stream);
BFPN_API bfpn_status bfpn_energy_half_f64(
    bfpn_workspace* ws, double* F, double* f,
    const double* S, const double* sigma, const double* T,
    double h, double dg, double density, int Ng, int nyz, int nang,
    const double* Q1, const double* Q2, int straggling, void* stream);

/* Catastrophic secondary source Q1 from the primary state.
 * Replaces secondarySourceFromPrimaryKernel (src/energy_solver.cu:287-333):
 *   Q1[g,s,m] = dg * sum_{g' in [beg_g,end_g)} sum_m'
 *               (K1[g,g'] max(Fp[g',s,m'],0) + K2[g,g'] fp[g',s,m']/3) Kv[g',m',m]
 * ker_e1/ker_e2 are Ng*Ng row-major FP64, ker_v is Ng*nang*nang FP64 [g'][in][out],
 * ker_beg/ker_end are the per-row nonzero column spans (Ng ints). */
BFPN_API bfpn_status bfpn_secondary_source_f32(
    bfpn_workspace* ws, const float* Fp, const float* fp, float* Q1,
    const double* ker_e1, const double* ker_e2, const double* ker_v,
    const int* ker_beg, const int* ker_end,
    int Ng, int nyz, int nang, double dg, void* stream);
BFPN_API bfpn_status bfpn_secondary_source_f64(
    bfpn_workspace* ws, const double* Fp, const double* fp, double* Q1,
    const double* ker_e1, const double* ker_e2, const double* ker_v,
    const int* ker_beg, const int* ker_end,
    int Ng, int nyz, int nang, double dg, void* stream);
/* Builds the source operands of these tables into the workspace (FP32: the
 * split-BF16 transposed Kv with its block-sparsity lists and the banded energy
 * matrices); later bfpn_secondary_source_* calls with the same table pointers,
 * Ng, nang and dg reuse them.  Call it again after changing table contents in
 * place.  Synchronizes `stream`.  Stands for the one-time kernel upload of
 * BFPnSolver::initialize (src/bfpm_solver.cu:1638-1813), which the
 * reference's per-step solveSecondaryFromPrimary relies on. */
BFPN_API bfpn_status bfpn_secondary_source_prepare_f32(
    bfpn_workspace* ws, const double* ker_e1, const double* ker_e2, const double* ker_v,
    const int* ker_beg, const int* ker_end, int Ng, int nang, double dg, void* stream);
BFPN_API bfpn_status bfpn_secondary_source_prepare_f64(
    bfpn_workspace* ws, const double* ker_e1, const double* ker_e2, const double* ker_v,
    const int* ker_beg, const int* ker_end, int Ng, int nang, double dg, void* stream);

/* Explicit MUSCL predictor-corrector transport over h, in place.
 * Replaces SpatialTransportSolver::solveEq18
 * (src/spatial_transport.cuh:41-44, src/spatial_transport.cu:23-161, 202-262).
 * omend holds nang (omega_y, omega_z) pairs.  preserve_sign = 0 clips to >= 0. */
BFPN_API bfpn_status bfpn_transport_half_f32(
    bfpn_workspace* ws, float* F, const double* omend,
    int Ny, int Nz, int Ng, int nang, double dy, double dz, double h,
    int preserve_sign, void* stream);
BFPN_API bfpn_status bfpn_transport_half_f64(
    bfpn_workspace* ws, double* F, const double* omend,
    int Ny, int Nz, int Ng, int nang, double dy, double dz, double h,
    int preserve_sign, void* stream);

/* Crank-Nicolson angular diffusion over dt on the zero-boundary DST-I basis,
 * in place.  Replaces AngleDiffusionSolver::solveEq19
 * (src/angle_diffusion.cuh:52-54, src/angle_diffusion.cu:7-34, 254-483).
 * sig_tr has Ng entries (host pointer: the CN factors are built on the host). */
BFPN_API bfpn_status bfpn_angle_full_f32(
    bfpn_workspace* ws, float* F, const double* sig_tr_host,
    int nyz, int Ng, int Nmu, int Nom, double du, double dv, double dt,
    double density, void* stream);
BFPN_API bfpn_status bfpn_angle_full_f64(
    bfpn_workspace* ws, double* F, const double* sig_tr_host,
    int nyz, int Ng, int Nmu, int Nom, double du, double dv, double dt,
    double density, void* stream);

/* Eq.24 integrated depth dose of (F+F1, f+f1).  Synchronizes `stream` and
 * returns the value in *idd (host).  Replaces computeIntegratedDepthDose + integratedDepthDoseKernel
 * + reduceDevicePartials (src/bfpm_solver.cu:229-309, 2132-2164, 3473-3502).
 * S/sigma/en are device pointers (Ng+1, Ng, Ng+1). F1/f1 may be NULL. */
BFPN_API bfpn_status bfpn_idd_f32(
    bfpn_workspace* ws, const float* F, const float* f, const float* F1, const float* f1,
    const double* S, const double* sigma, const double* en,
    int Ng, int Ny, int Nz, int nang, double dg, double du, double dv,
    double dy, double dz, double density, int trapezoidal_yz,
    int dose_catastrophic_loss, double* idd, void* stream);
BFPN_API bfpn_status bfpn_idd_f64(
    bfpn_workspace* ws, const double* F, const double* f, const double* F1, const double* f1,
    const double* S, const double* sigma, const double* en,
    int Ng, int Ny, int Nz, int nang, double dg, double du, double dv,
    double dy, double dz, double density, int trapezoidal_yz,
    int dose_catastrophic_loss, double* idd, void* stream);

/* bfpn_idd_f32 without a workspace and without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  Its scratch (dose coefficients, y/z weights and
 * per-block partial sums) is a fixed 2 MiB device array in the library, one
 * per GPU, shared by every call on that GPU: calls that overlap on one GPU
 * race and must be ordered by the caller.  Grids whose scratch exceeds 2 MiB
 * (Ng*16 + (Ny+1)(Nz+1)*8 + (Ny+1)(Nz+1)*nang/256*8 bytes, each part rounded
 * up to 256) are rejected with BFPN_E_INVALID.  Same arguments, result and
 * synchronization as bfpn_idd_f32. */
BFPN_API bfpn_status bfpn_idd_f32_noalloc(
    const float* F, const float* f, const float* F1, const float* f1,
    const double* S, const double* sigma, const double* en,
    int Ng, int Ny, int Nz, int nang, double dg, double du, double dv,
    double dy, double dz, double density, int trapezoidal_yz,
    int dose_catastrophic_loss, double* idd, void* stream);

/* ---------------------------------------------------------------------------
 * Context-level API: one beam, state resident in HBM.
 * Replaces BFPnSolver(GridParams, PhysicsParams), initialize(path) and
 * solve(tFinal) (src/bfpm_solver.cuh:166-178, src/bfpm_solver.cu:1171-1202,
 * 1435-1568, 2166-2267).
 * ------------------------------------------------------------------------- */

/* Builds tables on the host (FP64), allocates the state on the current
 * device, and initializes the Eq.23 beam on the device.  Fails with
 * BFPN_E_NOMEM when the estimated footprint exceeds free memory
 * (preflightMemory, src/bfpm_solver.cu:1321-1433). */
BFPN_API bfpn_status bfpn_create(const bfpn_grid* grid, const bfpn_phys* phys,
                                 const char* data_path, int precision,
                                 bfpn_ctx** out);

/* One layer of a medium stratified in depth: the material tables
 * (material_name as in bfpn_phys, data_path as in bfpn_create) and density
 * (g/cm^3) serve the depth steps whose midpoint (n + 1/2) dt lies below x_end
 * (cm); the last layer extends to any depth. */
typedef struct bfpn_layer {
    double x_end;
    const char* material;
    double density;
    const char* data_path;
} bfpn_layer;

/* bfpn_create for a layered medium.  No counterpart in the reference, whose
 * solver is homogeneous (the material-map item of TODO.md:114-118 and the
 * validation gap of README.md:30-37).  Each layer builds its own tables from
 * `phys` with the layer's material and density (both `density` and
 * `energy_density`); the Eq.23 beam comes from `phys`.  Depth steps use the
 * tables of their layer: the energy sweeps, the angular-diffusion factors,
 * the secondary source and the IDD scale 1/density.  Layer depths must
 * increase.  bfpn_get_table returns the first layer's tables. */
BFPN_API bfpn_status bfpn_create_layered(const bfpn_grid* grid, const bfpn_phys* phys,
                                         const bfpn_layer* layers, int nlayers,
                                         int precision, bfpn_ctx** out);

/* bfpn_create for a material map: per depth slab, the material of every
 * y-z node.  No counterpart in the reference (its material-map item,
 * TODO.md:114-118, is open).  materials[k] (x_end unused) builds a table set
 * from `phys` with that material and density, as a layer does.  Slab j covers
 * the depth steps whose midpoint lies below slab_end[j] (the last slab: any
 * depth); map[j (Ny+1)(Nz+1) + s] is the material index of node
 * s = iz (Ny+1) + iy in slab j.  The energy sweeps, angular diffusion,
 * secondary source and dose of a node use its material's tables and density;
 * transport is material-independent.  Streams (bfpn_set_overlap) and group
 * skipping (bfpn_set_group_skip) work as for bfpn_create; the legacy energy
 * model, lite memory, energy moments and spot planes are not available. */
BFPN_API bfpn_status bfpn_create_mapped(const bfpn_grid* grid, const bfpn_phys* phys,
                                        const bfpn_layer* materials, int nmat,
                                        const double* slab_end, int nslabs,
                                        const unsigned char* map, int precision, bfpn_ctx** ctx);
BFPN_API void bfpn_destroy(bfpn_ctx* ctx);

/* Derived deltas and array sizes. */
BFPN_API bfpn_status bfpn_get_deltas(const bfpn_ctx* ctx, bfpn_deltas* out);
BFPN_API size_t bfpn_cells(const bfpn_ctx* ctx);          /* Ng*nyz*nang */
BFPN_API size_t bfpn_device_bytes(const bfpn_ctx* ctx);   /* bytes held  */
/* Number of depth steps of a run to t_final: ceil(t_final/dt - 1e-12)
 * (src/bfpm_solver.cu:2178). */
BFPN_API int bfpn_num_steps(const bfpn_ctx* ctx, double t_final);

/* Host tables the context was built with (FP64 copies), for parity tests.
 * which: 0 S_s (Ng+1), 1 sigma_c (Ng), 2 T_c (Ng), 3 sig_trg (Ng), 4 en (Ng+1),
 * 5 ker_e1 (Ng*Ng), 6 ker_e2 (Ng*Ng), 7 ker_v (Ng*nang*nang), 8 omend (2*nang),
 * 9 mu (3*Ng), 10 ker_beg (Ng, as doubles), 11 ker_end (Ng, as doubles). */
BFPN_API bfpn_status bfpn_get_table(const bfpn_ctx* ctx, int which,
                                    double* out, size_t count);

/* Advance `nsteps` Strang depth steps (stepPrimary, stepSecondary when
 * step > 0, IDD and proxy; src/bfpm_solver.cu:2189-2255, 3238-3394).
 * Stream-ordered on the context's stream; returns without synchronizing. */
BFPN_API bfpn_status bfpn_advance(bfpn_ctx* ctx, int nsteps);

/* Waits for the context's stream, then copies the IDD and proxy history
 * (one value per completed step, x = (i+1)*dt) to the host. */
BFPN_API bfpn_status bfpn_history(bfpn_ctx* ctx, double* idd, double* proxy,
                                  int capacity, int* nsteps);

/* Runs a whole solve to t_final and writes the history (BFPnSolver::solve). */
BFPN_API bfpn_status bfpn_run(bfpn_ctx* ctx, double t_final, double* idd,
                              double* proxy, int capacity, int* nsteps);

/* Copies one state array to/from the host as FP64.
 * which: 0 = psi1 primary (F), 1 = psi2 primary (f_F), 2 = psi1 secondary (F1),
 * 3 = psi2 secondary (f_F1). */
BFPN_API bfpn_status bfpn_read_state(bfpn_ctx* ctx, int which, double* host,
                                     size_t count);
BFPN_API bfpn_status bfpn_write_state(bfpn_ctx* ctx, int which,
                                      const double* host, size_t count);

/* Paper Eq.25 energy moments of the current state: moments[2g] = du dv dy dz
 * sum over lanes of (F + F1) at group g, moments[2g+1] the same for (f + f1);
 * capacity >= 2*Ng.  Replaces BFPnSolver::computeEnergyMoments
 * (src/bfpm_solver.cu:3504-3526, energyMomentKernel 311-360). */
BFPN_API bfpn_status bfpn_energy_moments(bfpn_ctx* ctx, double* moments, int capacity);

/* Y-Z dose plane of the current state: per transverse node s = iz*(Ny+1)+iy,
 * the Eq.24 cell dose of psi^1 = F + F1 and psi^2 = f + f1 summed over
 * groups and directions, times du*dv/density; count = (Ny+1)*(Nz+1).
 * Replaces BFPnSolver::computeSpotDosePlane (src/bfpm_solver.cu:3560-3585,
 * spotDosePlaneKernel 433-475). */
BFPN_API bfpn_status bfpn_spot_plane(bfpn_ctx* ctx, double* plane, size_t count);

/* The HIP stream (hipStream_t) on which the context's issued work ends, so
 * callers can order their own work: work enqueued there runs after every
 * step issued so far, and steps issued later start after it. */
BFPN_API void* bfpn_stream(bfpn_ctx* ctx);

/* Overlap (on by default): with the secondary enabled, the source, the
 * secondary march and the IDD reduction of step n run on a second stream,
 * concurrently with the primary march of step n+1.  Off: every operator in
 * issue order on one stream (the reference's sequence; per-kernel timings
 * without contention).  Results are identical either way.  No counterpart in
 * the reference, whose operators synchronise the host after every call. */
BFPN_API bfpn_status bfpn_set_overlap(bfpn_ctx* ctx, int on);

/* Active-group skipping (on by default).  Above the highest nonzero energy
 * group of a species every group of its state is exactly zero (energy only
 * decreases, straggling lifts it by at most one group per half step, the
 * secondary source emits at or below its source group, transport and
 * angular diffusion keep groups apart).  Each energy sweep measures that top
 * group on the device, and the operators of the next sweeps skip the groups
 * above it, whose results would be exact zeros: results are bit-identical
 * with skipping off.  The reference's streaming mode skips zero chunks the
 * same way from a conservative bound (src/bfpm_solver.cu:2325-2342,
 * --no-zero-chunk-skip); here the bound is exact and also falls as the beam
 * slows.  Call between bfpn_advance calls. */
BFPN_API bfpn_status bfpn_set_group_skip(bfpn_ctx* ctx, int on);
/* The current top groups of the primary and secondary state (-1: all zero;
 * Ng-1 when skipping is off).  Synchronizes. */
BFPN_API bfpn_status bfpn_active_groups(bfpn_ctx* ctx, int* primary_top, int* secondary_top);

/* Per-phase device timing of the bfpn_advance calls made while profiling is
 * on (replaces StepProfile, src/bfpm_solver.cuh:64-75).  Each operator launch
 * is bracketed by HIP events on the context's stream.  Phases: 0 energy,
 * 1 transport, 2 angle, 3 source, 4 dose/history.  For each phase: summed
 * seconds, number of operator calls, and the algorithmic HBM bytes those
 * calls must move (each state array read once and written once per call,
 * plus the source's intermediate arrays).  Turning profiling on resets them. */
BFPN_API bfpn_status bfpn_set_profiling(bfpn_ctx* ctx, int on);
BFPN_API bfpn_status bfpn_phase_stats(bfpn_ctx* ctx, double* seconds5, int* calls5,
                                      double* bytes5);

/* Stopping power table S on the Ng+1 group edges of [1, 1+Lg] MeV for a
 * material ("", water, bone, air) -- computeStoppingPower,
 * src/bfpm_solver.cu:1570-1597.  Host only; used e.g. for CSDA ranges. */
BFPN_API bfpn_status bfpn_stopping_power(const char* material, int Ng, double Lg, double* S);

/* ---------------------------------------------------------------------------
 * Serving: continuous batching of beamlet solves (no counterpart in the
 * reference, which runs one beam per process).  A scheduler thread admits
 * queued requests as slots and HBM budget free up, keeps each admitted beam
 * resident on its own stream, interleaves slices of depth steps of all
 * running beams, and shares physics tables between beams with equal grid,
 * step, material and flags.  A beam's IDD does not depend on the other
 * beams in flight.
 * ------------------------------------------------------------------------- */
typedef struct bfpn_server bfpn_server;

typedef struct bfpn_server_opts {
    size_t hbm_budget_bytes;  /* 0 = 90 % of the free device memory at creation */
    int max_active;           /* beams resident and running at once            */
    int slice_steps;          /* depth steps enqueued per beam per turn        */
    int precision;            /* BFPN_FP32 or BFPN_FP64                        */
    size_t fill_cells;        /* phase-space cells kept in flight, shortest
                                 remaining work first (0 = 4e7: a beam that
                                 fills the GPU runs alone)                    */
} bfpn_server_opts;

typedef struct bfpn_server_stats {
    uint64_t submitted, completed, failed;
    int active, queued, peak_active, cached_tables;
    size_t bytes_in_use;
} bfpn_server_stats;

BFPN_API bfpn_status bfpn_server_create(const bfpn_server_opts* opts, bfpn_server** out);
/* Queues one beam solve to t_final (same grid/phys/data meaning as bfpn_create);
 * returns at once with its id. Thread safe. */
BFPN_API bfpn_status bfpn_server_submit(bfpn_server* s, const bfpn_grid* grid,
                                        const bfpn_phys* phys, const char* data_path,
                                        double t_final, uint64_t* id);
/* The same for a medium layered in depth (layers as in bfpn_create_layered).
 * No counterpart in the reference.  Each layer's tables are cached under
 * their own key, so a layer of the same material, density and tables as a
 * homogeneous request shares that request's tables. */
BFPN_API bfpn_status bfpn_server_submit_layered(bfpn_server* s, const bfpn_grid* grid,
                                                const bfpn_phys* phys, const bfpn_layer* layers,
                                                int nlayers, double t_final, uint64_t* id);
/* The same under a material map (materials, slabs and map as in
 * bfpn_create_mapped; no counterpart in the reference).  Each material's
 * tables are cached as a layer's are. */
BFPN_API bfpn_status bfpn_server_submit_mapped(bfpn_server* s, const bfpn_grid* grid,
                                               const bfpn_phys* phys, const bfpn_layer* materials,
                                               int nmat, const double* slab_end, int nslabs,
                                               const unsigned char* map, double t_final,
                                               uint64_t* id);
/* 0 queued, 1 running, 2 done, 3 failed, -1 unknown (or already collected). */
BFPN_API int bfpn_server_poll(bfpn_server* s, uint64_t id);
/* Blocks until request `id` finished (timeout_s <= 0: no limit), copies its IDD
 * history and releases the result.  times3 = {queued s, running s, total s}. */
BFPN_API bfpn_status bfpn_server_wait(bfpn_server* s, uint64_t id, double timeout_s,
                                      double* idd, int capacity, int* nsteps,
                                      double* times3);
BFPN_API bfpn_status bfpn_server_get_stats(bfpn_server* s, bfpn_server_stats* out);
/* The admission estimates the scheduler charges against the budget: device
 * bytes of one beam's own arrays, and of the physics tables beams with the
 * same grid, material and flags share.  Host only. */
BFPN_API bfpn_status bfpn_server_estimate(const bfpn_grid* grid, const bfpn_phys* phys,
                                          int precision, size_t* context_bytes,
                                          size_t* physics_bytes);
/* Finishes queued and running requests, then stops the scheduler. */
BFPN_API void bfpn_server_destroy(bfpn_server* s);

#ifdef __cplusplus
}
#endif

#endif /* BFPN_H */
