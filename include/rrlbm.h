/*
 * rrlbm.h -- C ABI of the MI355X (gfx950) recursively-regularised
 * lattice-Boltzmann hot path.
 *
 * The reference (RR27V33_code_DNS_LES/) exposes its hot path as CUDA kernels
 * on global FP64 device arrays, called in a fixed order by main()
 * (RR27V33_code_DNS_LES/LBMLES.h:58-81, main.cu:31-117).  This header offers
 *
 *   - one drop-in entry point per reference kernel, operating on FP64 fields in
 *     the reference layout that the context owns (populations SoA
 *     f[ip*N + idx], idx = (iz*LY + iy)*LX + ix, plus rho, ux, uy, uz,
 *     force_realx/y/z and tau_t), so that calling them in the main.cu order
 *     reproduces the reference step;
 *   - the fused path: one kernel per time step doing pull-streaming,
 *     bounce-back, macrovar, RRBC, macrovar_BN, WALE and the RR collision, on
 *     shifted populations stored in FP64, FP32 or per-direction-scaled FP16;
 *   - FP64 diagnostics (diag_flow sums and output_profile plane statistics).
 *
 * Conventions: every call returns an rrlbm_status (0 = success) and never
 * exits the process; rrlbm_last_error() gives the message of the last failure
 * of a context (or of rrlbm_create when ctx is NULL, per host thread).  Work is
 * enqueued on the context's own HIP stream; calls that return host data
 * synchronise that stream.  One context per host thread; the only state
 * shared between contexts is the process's autotuning cache (rrlbm_autotune,
 * mutex-protected).  Sizes are counts of doubles unless stated.
 *
 * Velocity ordering: NPOP = 27 uses the channel solver's order
 * (RR27V33_code_DNS_LES/initial.cu:67-69); NPOP = 19 uses the Taylor-Green
 * D3Q19 order (Taylor_Green/D3Q19RR.f90:75-77).  rrlbm_lattice() returns them.
 */
#ifndef RRLBM_H
#define RRLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RRLBM_ABI_VERSION 1

typedef enum rrlbm_status {
    RRLBM_OK = 0,
    RRLBM_E_INVALID = 1,     /* bad argument or parameter combination */
    RRLBM_E_HIP = 2,         /* HIP runtime error (message has the HIP error string) */
    RRLBM_E_STATE = 3,       /* call out of order (e.g. fused step before pack) */
    RRLBM_E_NUMERIC = 4,     /* non-finite values detected by a diagnostics call */
    RRLBM_E_IO = 5,          /* file open / read / write failure */
    RRLBM_E_UNSUPPORTED = 6  /* valid request the library does not provide */
} rrlbm_status;

/* collision operator: the RR operator of the channel solver (RR27V33_code_DNS_LES/
 * collision.cu) or the BGK baseline of the reference's BGK solver (BGK/collision.cu:
 * incompressible second-order equilibrium, u = j + F/2 without division by rho,
 * BGK/macrovar.cu:27; D3Q27, no LES, no RR wall regularisation).  The library's
 * reference-layout fields keep their convention under BGK: f is unshifted and rho
 * is 1 + the BGK solver's density deviation (that solver stores f - w and rho - 1). */
typedef enum rrlbm_collision { RRLBM_COLLISION_RR = 0, RRLBM_COLLISION_BGK = 1 } rrlbm_collision;

/* storage of the fused path's populations */
typedef enum rrlbm_storage { RRLBM_STORE_F64 = 0, RRLBM_STORE_F32 = 1, RRLBM_STORE_F16 = 2 } rrlbm_storage;

/* rounding of the FP16 store.  Round-to-nearest loses every per-step change
 * smaller than half an FP16 step (a slowly evolving flow stalls: a laminar
 * channel stops accelerating under its body force); stochastic rounding keeps
 * such changes in expectation.  The dither is a fixed function of (cell,
 * population, step, instance), so runs are reproducible. */
typedef enum rrlbm_f16_rounding { RRLBM_F16_STOCHASTIC = 0, RRLBM_F16_NEAREST = 1 } rrlbm_f16_rounding;

/* reference-layout FP64 fields owned by the context (RR27V33_code_DNS_LES/globals.cu:67-71) */
typedef enum rrlbm_field {
    RRLBM_FIELD_F = 0,      /* d_f, NPOP*N */
    RRLBM_FIELD_F_TEMP = 1, /* d_f_temp, NPOP*N */
    RRLBM_FIELD_RHO = 2,    /* d_rho, N */
    RRLBM_FIELD_UX = 3,
    RRLBM_FIELD_UY = 4,
    RRLBM_FIELD_UZ = 5,
    RRLBM_FIELD_FORCEX = 6, /* d_force_realx, N */
    RRLBM_FIELD_FORCEY = 7,
    RRLBM_FIELD_FORCEZ = 8,
    RRLBM_FIELD_TAUT = 9    /* d_taut, N */
} rrlbm_field;

/* Run-time counterpart of parameters.h (RR27V33_code_DNS_LES/parameters.h:5-24)
 * and the physics globals (globals.cu:57-63). */
typedef struct rrlbm_params {
    int32_t lx, ly, lz;         /* LX, LY, LZ */
    int32_t npop;               /* NPOP: 27 (D3Q27) or 19 (D3Q19) */
    int32_t use_les;            /* USE_LES: 0 DNS, 1 WALE (D3Q27 only) */
    int32_t les_lagged;         /* fused path with LES: 1 = WALE from the previous step's
                                   velocity (one launch per step); 0 = reference order
                                   (a velocity pass, then the step: two launches) */
    int32_t storage;            /* rrlbm_storage of the fused populations */
    int32_t wall_x, wall_y, wall_z; /* 1: half-way bounce-back walls on both faces of that
                                   axis (channel: wall_y); 0: periodic */
    int32_t wall_rr;            /* 1: RR regularisation of wall-incoming populations (RRBC);
                                   2: the finite-difference reconstruction FDBC in its place
                                   (D3Q27, y walls only, ly >= 4; the fused path without LES) */
    int32_t lid_face;           /* -1 none, else face 2*axis + side (side 0 low, 1 high)
                                   moving with lid_u (cavity) */
    double tau;                 /* d_tau = 3 visc + 1/2 */
    double force[3];            /* uniform body force of the fused path (d_force_in_x, 0, 0) */
    double cw;                  /* WALE constant (LES.cu:12) */
    double lid_u[3];            /* lid velocity */
    double f16_scale;           /* FP16 storage: (f_i - w_i) * f16_scale / w_i is stored;
                                   0 selects 4096 */
    int32_t xcd_remap;          /* order of the blocks of cells over the workgroup IDs, which the
                                   dispatcher deals round-robin to the 8 XCDs: 0 linear;
                                   1 one contiguous slab per XCD; C (2..65536) runs of C
                                   consecutive blocks per XCD; -1 = library default (32 on grids
                                   of at least 2048 blocks of 256 cells, else 0) */
    int32_t cells_per_thread;   /* fused D3Q27 step without LES: 1 or 2 cells per thread
                                   (64 cells apart); 0 = the tuned / default choice */
    int32_t f16_rounding;       /* rrlbm_f16_rounding of the FP16 store (default stochastic) */
    int32_t pop_group_rows;     /* layout of the fused populations: groups of this many whole
                                   rows (x lines) store their NPOP populations together;
                                   -1 = one group (structure of arrays), 0 = library default
                                   (groups of about 6144 cells when LX is a multiple of 64,
                                   else one group); rrlbm_get_params reports the value used */
    int32_t pop_pad_bytes;      /* bytes of padding after each population segment (multiple of 16,
                                   at most 1 MiB); -1 = library default */
    int32_t persistent_wg_per_cu; /* fused launches: 0 = one workgroup per block of cells;
                                   k > 0 = a persistent grid of k workgroups per compute unit
                                   striding over the blocks */
    int32_t collision_model;    /* rrlbm_collision (default RR) */
    int32_t rr_order;           /* order of the RR expansion: 3 (0 = default) as the reference
                                   runs it, or 6 with the fourth- to sixth-order Hermite terms the
                                   reference keeps commented out (collision.cu:48-49,62-67,95-104,
                                   112-116; BC.cu:37-38,50-54,79-88,97-101); 6: D3Q27, RR, no LES */
    int32_t steps_per_pass;     /* fused steps per launch: 0 or 1 = one; 2 = two steps in one pass
                                   over the populations (the first step's state stays on chip), for
                                   the D3Q27 RR (orders 3 and 6) and BGK steps without LES or FDBC
                                   in FP32 or FP64 storage with lx a multiple of 64 and ly of 4;
                                   an odd step count and forcing_GP steps take single launches */
    int32_t pass_planes;        /* steps_per_pass 2: z planes per workgroup's chunk (0 = library
                                   default, 16) */
} rrlbm_params;

typedef struct rrlbm_ctx rrlbm_ctx;

/* Diagnostics (FP64).  The sums follow diag_flow (output.cu:24-35). */
typedef struct rrlbm_diag {
    double mass_excess;        /* sum (rho - 1) */
    double u_sum[3];           /* sum u, v, w (diag_flow's umean*N, vmean*N, wmean*N) */
    double momentum[3];        /* sum rho u */
    double kinetic_energy;     /* sum rho |u|^2 / 2 */
    double nonfinite_cells;    /* cells with a non-finite rho or kinetic energy */
    double ncells;             /* N */
} rrlbm_diag;

/* ---------------------------------------------------------------- setup */
/* Library version string. */
const char* rrlbm_version(void);
/* Defaults of parameters.h / globals.cu: 64^3, D3Q27, DNS, channel walls in y,
 * RRBC on, Re_tau = 180 physics (globals.cu:57-63). */
int rrlbm_params_default(rrlbm_params* p);
/* Velocity set of npop (27 or 19): c[3*i+k], w[i], opp[i]. */
int rrlbm_lattice(int npop, int32_t* c, double* w, int32_t* opp);
/* Number of visible HIP devices. */
int rrlbm_device_count(int* n);
/* Allocates the reference fields and the fused buffers on the current device
 * (replaces initialize()'s allocation and constant upload, initial.cu:4-178). */
int rrlbm_create(const rrlbm_params* p, rrlbm_ctx** out);
/* Frees everything (main.cu:133-150). */
int rrlbm_destroy(rrlbm_ctx* ctx);
/* Message of the last failure of ctx (ctx NULL: last rrlbm_create failure of this thread). */
const char* rrlbm_last_error(const rrlbm_ctx* ctx);
/* Parameters the context was created with. */
int rrlbm_get_params(const rrlbm_ctx* ctx, rrlbm_params* out);
/* Waits for the context's stream. */
int rrlbm_synchronize(rrlbm_ctx* ctx);
/* Device bytes held by the context. */
int rrlbm_memory_bytes(const rrlbm_ctx* ctx, size_t* bytes);

/* ------------------------------------------------ reference-layout fields */
/* Host <-> device copies of one FP64 field (cudaMemcpy of initial.cu:237-260,
 * output.cu:19-22,167-171,234).  count must equal the field's size. */
int rrlbm_upload(rrlbm_ctx* ctx, rrlbm_field field, const double* host, size_t count);
int rrlbm_download(rrlbm_ctx* ctx, rrlbm_field field, double* host, size_t count);

/* ------------------------------------------- drop-ins for reference kernels */
/* init_arrays, initial.cu:268-325 (ustar = u0, ystar = visc/u0 of globals.cu:59-61;
 * the force is params.force[0]). */
int rrlbm_init_arrays(rrlbm_ctx* ctx, double ustar, double ystar);
/* forcing, macrovar.cu:63-74 (writes params.force into the force fields). */
int rrlbm_forcing(rrlbm_ctx* ctx);
/* forcing_GP, macrovar.cu:76-114 (force_in_x = params.force[0]). */
int rrlbm_forcing_gp(rrlbm_ctx* ctx, int istep, double rand_phase);
/* LES_WALE, LES.cu:4-98 (tau_t from the velocity fields). */
int rrlbm_les_wale(rrlbm_ctx* ctx);
/* collision_RR, collision.cu:4-119 (in place; uses the force fields and, with
 * use_les, tau_t). */
int rrlbm_collision_rr(rrlbm_ctx* ctx);
/* streaming, streaming.cu:3-24 (push f -> f_temp, periodic). */
int rrlbm_streaming(rrlbm_ctx* ctx);
/* BB, streaming.cu:26-55 (half-way bounce-back into f_temp at walled faces). */
int rrlbm_bb(rrlbm_ctx* ctx);
/* The D2D copy f_temp -> f of main.cu:64, done as a buffer swap: afterwards
 * f holds the streamed populations and f_temp the post-collision ones. */
int rrlbm_swap_f(rrlbm_ctx* ctx);
/* macrovar, macrovar.cu:4-31 (BGK model: BGK/macrovar.cu:4-31, u = j + F/2). */
int rrlbm_macrovar(rrlbm_ctx* ctx);
/* collision_BGK, BGK/collision.cu:3-33 (in place; needs collision_model = BGK). */
int rrlbm_collision_bgk(rrlbm_ctx* ctx);
/* RRBC, BC.cu:4-126 (needs rrlbm_swap_f to have run: re-imposed bounce-back
 * values at multi-face cells come from f_temp). */
int rrlbm_rrbc(rrlbm_ctx* ctx);
/* FDBC, BC.cu:129-309: at the y-wall rows, velocity extrapolated from the two
 * next rows, strain rate by finite differences, density from the known
 * populations, all populations rebuilt as feq + fneq - Fbar feq (the
 * alternative to RRBC at main.cu:76-77).  The wall velocities are all
 * extrapolated before any difference is taken (the reference's single kernel
 * races there).  D3Q27, y walls only, ly >= 4, third order. */
int rrlbm_fdbc(rrlbm_ctx* ctx);
/* macrovar_BN, macrovar.cu:33-62. */
int rrlbm_macrovar_bn(rrlbm_ctx* ctx);
/* nsteps of main.cu:50-84 with the reference kernels: [LES_WALE] ->
 * collision_RR -> streaming -> BB -> copy -> macrovar -> [RRBC (FDBC with wall_rr 2) -> macrovar_BN]
 * (BGK model: BGK/main.cu:49-70, collision_BGK -> streaming -> BB -> copy ->
 * macrovar).  The force fields are used as they are. */
int rrlbm_step_reference(rrlbm_ctx* ctx, int nsteps);

/* ------------------------------------------------------- initialisation */
/* Configuration initialisers (no reference counterpart: the configurations'
 * seeded fields).  They write rho and u and set f to the RR equilibrium. */
int rrlbm_init_rest(rrlbm_ctx* ctx);
int rrlbm_init_taylor_green_2d(rrlbm_ctx* ctx, double u0);
int rrlbm_init_taylor_green_3d(rrlbm_ctx* ctx, double u0);
/* Channel: u_x = u_tau U+(y+) (viscous / ln law), streaks amp_streak*U(y), noise
 * amp_noise*U_bulk on every component, seeded counter-based RNG. */
int rrlbm_init_channel(rrlbm_ctx* ctx, double u_tau, double visc, uint64_t seed, double amp_streak,
                       double amp_noise);
/* f = RR equilibrium of the current rho and u fields. */
int rrlbm_init_equilibrium(rrlbm_ctx* ctx);

/* ------------------------------------------------------------ fused path */
/* Collide the reference-layout state (the next step's collision_RR, using
 * params.force -- or the force fields when rrlbm_forcing_gp wrote them last --
 * and, with use_les, tau_t from LES_WALE) into the fused buffer. */
int rrlbm_fused_pack(rrlbm_ctx* ctx);
/* nsteps fused time steps under the uniform params.force. */
int rrlbm_fused_step(rrlbm_ctx* ctx, int nsteps);
/* One fused time step whose collision is step istep of main.cu's loop during
 * its forcing_GP phase (main.cu:33-46, macrovar.cu:76-114): the force is
 * computed per cell in the kernel (no force fields are read or written); the
 * step the launch completes keeps the force its collision had.  D3Q27 RR
 * channel, params.force = (force_in_x, 0, 0).  The loop of main.cu becomes
 *   forcing_gp(istep0) + fused_pack, then fused_step_gp(istep) while
 *   istep < npforcing, then fused_step. */
int rrlbm_fused_step_gp(rrlbm_ctx* ctx, int istep, double rand_phase);
/* As rrlbm_fused_step, timing the launches with HIP events on the context's
 * stream; *ms receives the elapsed milliseconds. */
int rrlbm_fused_step_timed(rrlbm_ctx* ctx, int nsteps, double* ms);
/* Capture nsteps fused steps into a HIP graph (replaces a previous capture);
 * the reference launches every kernel from the host each step (main.cu:31-117).
 * Replays must start from the buffer the capture started from: when the
 * captured launches swap the buffers an odd number of times (an odd nsteps; with
 * steps_per_pass 2 one swap per two steps, so nsteps = 2 mod 4 as well) every
 * second replay is refused (RRLBM_E_STATE; capture again).  With
 * FP16 stochastic rounding the captured steps read their step number from a
 * device counter of the context that each replay advances, so replays give
 * the state of as many eager steps, dither included.  A change of population
 * layout (rrlbm_autotune) drops the captured graph. */
int rrlbm_graph_capture(rrlbm_ctx* ctx, int nsteps);
/* Replay the captured graph `times` times (times * nsteps steps) under the
 * uniform force; RRLBM_E_STATE without a packed state, or when the state's
 * last collision took forcing_GP (run rrlbm_fused_step once first). */
int rrlbm_graph_launch(rrlbm_ctx* ctx, int times);
/* As rrlbm_graph_launch, timed with HIP events on the context's stream (ms). */
int rrlbm_graph_launch_timed(rrlbm_ctx* ctx, int times, double* ms);
/* Finish the current step into the reference fields: streaming, BB, macrovar,
 * RRBC and macrovar_BN, writing f, rho, u (FP64). */
int rrlbm_fused_unpack(rrlbm_ctx* ctx);
/* FP64 diagnostics of the fused state (end-of-step rho, u; no state change). */
int rrlbm_fused_diag(rrlbm_ctx* ctx, rrlbm_diag* out);

/* ------------------------------------------------------------ autotuning */
/* Launch shape and population layout chosen for a context. */
typedef struct rrlbm_tuning {
    int32_t pop_group_rows;       /* as in rrlbm_params (value used) */
    int32_t pop_pad_bytes;
    int32_t cells_per_thread;
    int32_t xcd_remap;
    int32_t persistent_wg_per_cu;
    int32_t candidates;           /* configurations timed; 0 = taken from this process's cache */
    double ms_per_step;           /* fused step time of the choice in the confirmation trial */
    double ms_per_step_default;   /* the library default's time in the same trial */
    int32_t steps_per_pass;       /* as in rrlbm_params (value used) */
    int32_t pass_planes;
} rrlbm_tuning;
/* Times ctx's fused step over candidate population layouts (rows per group,
 * padding) and launch shapes (1 or 2 cells per thread, block order over the
 * XCDs: linear, slabs, runs of 4, 16, 32 or 64 blocks; one workgroup per block or a
 * persistent grid of 4 or 8 per compute unit; where the configuration allows, two
 * steps per launch with z chunks of 8, 12, 16, 24 or 32 planes), best of three interleaved
 * rounds of trial_steps steps each; then times the fastest against the
 * library default in rounds of 3 * trial_steps and keeps the default unless
 * the fastest still beats it.  The choice applies to ctx and is cached per
 * configuration (grid, lattice, storage, walls, LES, force components) for
 * later calls in this process.  It steps ctx's own fused buffers, so it runs
 * before rrlbm_fused_pack (RRLBM_E_STATE otherwise).  No reference counterpart:
 * the reference fixes its 8x8x4 block at compile time (parameters.h:22-24, main.cu:26-27). */
int rrlbm_autotune(rrlbm_ctx* ctx, int trial_steps, rrlbm_tuning* out);

/* ------------------------------------------------------------ diagnostics */
/* diag_flow sums (output.cu:3-66) over the reference fields. */
int rrlbm_diag_flow(rrlbm_ctx* ctx, rrlbm_diag* out);
/* Largest cell count of rrlbm_diag_flow_noalloc: its scratch array holds the
 * partial sums of this many cells (4.5 MiB of device memory per GPU). */
#define RRLBM_NOALLOC_MAX_CELLS (1 << 24)
/* rrlbm_diag_flow without a context and without allocating device memory: the
 * diag_flow sums of caller-owned FP64 device fields rho, ux, uy, uz of n cells,
 * by the kernels of rrlbm_diag_flow (identical results), enqueued on the HIP
 * stream `stream` (a hipStream_t; NULL: the null stream), which the call then
 * synchronises to fill *out.  The partial sums live in a fixed device array of
 * the library that each GPU holds until the process ends, so the call makes no
 * hipMalloc / hipFree; 1 <= n <= RRLBM_NOALLOC_MAX_CELLS (RRLBM_E_INVALID
 * otherwise).  All calls on one GPU share that array: calls that overlap in
 * time on one GPU (other host threads, other streams) race.  Errors are read
 * with rrlbm_last_error(NULL).  No counterpart in the reference: this is for
 * callers that cannot allocate device memory (allocation-free serving loops,
 * harnesses that forbid hipMalloc). */
int rrlbm_diag_flow_noalloc(const double* rho, const double* ux, const double* uy, const double* uz, int64_t n,
                            rrlbm_diag* out, void* stream);
/* output_profile statistics (output.cu:142-225) over the reference fields:
 * out[iy*10 + k] = <u>,<v>,<w>,<u'u'>,<v'v'>,<w'w'>,<u'v'>,<u'w'>,<v'w'>,<tau_t>, count = 10*LY. */
int rrlbm_plane_profiles(rrlbm_ctx* ctx, double* out, size_t count);

/* ---------------------------------------------------------------- files */
/* output_flow / output_f / restart (output.cu:68-140,227-258, initial.cu:184-265):
 * data/%09d.dat holds rho, ux, uy, uz (N doubles each); data/f%09d.dat holds
 * the NPOP*N populations, both raw FP64 in the reference layout. */
int rrlbm_write_flow(rrlbm_ctx* ctx, const char* dir, int istep);
int rrlbm_write_f(rrlbm_ctx* ctx, const char* dir, int istep);
int rrlbm_read_restart(rrlbm_ctx* ctx, const char* dir, int istep);

/* Frees the reference-layout FP64 fields of a context whose state lives in
 * the fused buffers (saves (2*NPOP + 8) * 8 bytes per cell); calls that need
 * those fields then return RRLBM_E_STATE.  No reference counterpart. */
int rrlbm_release_reference(rrlbm_ctx* ctx);

/* ------------------------------------------------------- batched ensembles */
/* Many periodic D3Q27 boxes of uneven size, relaxation time and velocity
 * stepped together (no reference counterpart: the reference runs one domain
 * per process).  One launch per step covers every instance: the launch walks a
 * work list of tiles (instance, block of 256 cells), one workgroup per tile or
 * a persistent grid (rrlbm_batch_set_grid / rrlbm_batch_autotune). */
typedef struct rrlbm_batch rrlbm_batch;
typedef struct rrlbm_instance {
    int32_t lx, ly, lz;
    int32_t reserved;
    double tau;   /* 3 nu + 1/2 of this instance */
    double u0;    /* initial Taylor-Green amplitude */
} rrlbm_instance;
/* keep_reference = 0 frees each instance's FP64 reference fields after packing. */
int rrlbm_batch_create(const rrlbm_instance* inst, int count, int storage, int keep_reference, rrlbm_batch** out);
int rrlbm_batch_destroy(rrlbm_batch* b);
const char* rrlbm_batch_last_error(const rrlbm_batch* b);
/* Total cells over all instances. */
int rrlbm_batch_cells(const rrlbm_batch* b, int64_t* cells);
/* Every instance: 3D Taylor-Green fields with its u0, RR equilibrium, packed. */
int rrlbm_batch_init_taylor_green_3d(rrlbm_batch* b);
/* nsteps, one batched launch per step. */
int rrlbm_batch_step(rrlbm_batch* b, int nsteps);
/* nsteps, one launch per instance per step, in order on one stream. */
int rrlbm_batch_step_sequential(rrlbm_batch* b, int nsteps);
/* Capture nsteps batched launches into a HIP graph (replaces a previous one).
 * With FP16 storage the captured steps read their step number from a device
 * counter that each replay advances: replays equal as many eager steps. */
int rrlbm_batch_graph_capture(rrlbm_batch* b, int nsteps);
/* Launch the captured graph `times` times (times * nsteps steps). */
int rrlbm_batch_graph_launch(rrlbm_batch* b, int times);
/* Time nsteps (mode 0 batched launches, 1 sequential launches) or `nsteps`
 * graph replays (mode 2) with HIP events on the batch stream. */
int rrlbm_batch_time(rrlbm_batch* b, int mode, int nsteps, double* ms);
/* FP64 diagnostics of one instance's current state. */
int rrlbm_batch_diag(rrlbm_batch* b, int instance, rrlbm_diag* out);
/* Grid of the batched launch: 0 = one workgroup per tile of the work list,
 * k > 0 = a persistent grid of k workgroups per compute unit walking it. */
int rrlbm_batch_set_grid(rrlbm_batch* b, int wg_per_cu);
/* Times trial steps for grids of 0, 2, 4, 8 and 16 workgroups per compute unit
 * and keeps the fastest; runs before rrlbm_batch_init_taylor_green_3d (it
 * steps the instances' fused buffers). */
int rrlbm_batch_autotune(rrlbm_batch* b, int trial_steps, int* wg_per_cu, double* ms_per_step);
/* Borrowed context of one instance (valid until rrlbm_batch_destroy), e.g. to
 * unpack and download its fields. */
int rrlbm_batch_context(rrlbm_batch* b, int instance, rrlbm_ctx** ctx);

#ifdef __cplusplus
}
#endif

#endif /* RRLBM_H */
