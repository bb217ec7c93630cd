/*
 * sharpclaw_amd: MI355X (gfx950) implementation of SharpClaw's semi-discrete
 * step -- WENO5 reconstruction, Riemann solve and SSP Runge-Kutta stage --
 * behind a plain C ABI.  Every array crossing this boundary is a host array
 * in the reference's Fortran layout: column-major, lower bound 1-mbc, one
 * contiguous plane per component (src/1d/qalloc1.f90:10-29,
 * src/2d/qalloc2.f90:11-30), i.e. q(1-mbc:nx+mbc[,1-mbc:ny+mbc], meqn).
 *
 * All entry points return 0 on success and a nonzero SC_ERR_* code on
 * failure; sc_last_error() then describes the failure.  The reference
 * reports the same conditions with write(6,*) + stop (src/flux1.f90:111-115,
 * src/1d/bc1.f90:34-38, src/main.f90:140-145).
 *
 * Threading: a solver handle is used by one host thread at a time; its work
 * is ordered on one HIP stream owned by the handle.
 */
#ifndef SHARPCLAW_AMD_H
#define SHARPCLAW_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define SC_OK 0
#define SC_ERR_ARG 1         /* invalid argument or option (reference: stop) */
#define SC_ERR_UNSUPPORTED 2 /* option the device path does not implement */
#define SC_ERR_HIP 3         /* HIP runtime failure */
#define SC_ERR_NOMEM 4       /* device allocation failed (qalloc STAT check) */
#define SC_ERR_STATE 5       /* call out of order (e.g. no q uploaded) */

/* physics: the Riemann solver / eigenvector plugin pair on the device.
 * Each id names the reference plugin it replaces:
 *   ADVECTION      apps/advection/1d/rp1.f + evec.f            (meqn=mwaves=1)
 *   ACOUSTICS      apps/acoustics/1d/homogeneous/rp1.f + evec.f (meqn=mwaves=2)
 *   EULER     1D   apps/euler/shockentropy/rp1eu.f + Euler evec (meqn=mwaves=3)
 *             2D   Roe rpn2 for Euler + direction-aware evec   (meqn=4, mwaves=3)
 *   SHALLOW_WATER 1D  .../1dRadialSymm-qwave/rp1sw.f + evec     (meqn=mwaves=2)
 *                 2D  apps/shallowWater/2dRadialDam/rpn2sw.f + evec (meqn=mwaves=3) */
#define SC_PHYS_ADVECTION 1
#define SC_PHYS_ACOUSTICS 2
#define SC_PHYS_EULER 3
#define SC_PHYS_SHALLOW_WATER 4

/* arithmetic of reconstruction + Riemann solve.  FP32 keeps float64 state,
 * float64 stage accumulation, dt/CFL selection and conservation sums. */
#define SC_FP64 0
#define SC_FP32 1

/* Everything module ClawParams (src/ClawParams.f90:4-14) and the physics
 * common blocks (/comrp/, /param/, /cparam/) carry into the hot path. */
typedef struct sc_params {
    int ndim;            /* 1 or 2 */
    int nx[2];           /* interior cells per direction (nx[1] ignored in 1D) */
    int meqn;
    int mwaves;
    int mbc;             /* must be 3 (reconstruct.f90:85,158 hard-code it) */
    int maux;            /* aux arrays per cell (sc_set_aux): 1D grids, 2D
                            advection's velocities, 2D capacity */
    int mcapa;           /* capacity column of aux, 1-based; 0 none (flux1.f90:56-61) */
    int lim_type;        /* 2 WENO5; 1 second-order TVD (single grids: 1D, 2D Euler
                            and shallow water) */
    int char_decomp;     /* lim_type 2: 0 weno5, 1 weno5_wave, 2 weno5_char,
                            3 weno5_trans; lim_type 1: 1 tvd2_wave, 2 tvd2_char
                            (reconstruct.f90:77-763, dispatch flux1.f90:65-117).
                            1 and 3, and lim_type 1, on single grids: 1D, and 2D
                            Euler / shallow water (x and y slice sweeps) */
    int tfluct_solver;   /* 0 in-cell problem through rp (flux1.f90:155-188); 1 total
                            fluctuation (flux1.f90:136-153), 1D advection/acoustics/
                            shallow water only: the applications with a tfluct */
    int time_integrator; /* 1 SSP22, 2 SSP42, 3 SSP33, 4 SSP(10,4) */
    int src_term;        /* source: 0 none (src1.f90/src2.f90: dq = 0); 1 the radial
                            shallow-water source of apps/shallowWater/2dRadialDam/
                            1dRadialSymm-qwave/src1.f90 (1D shallow water) */
    int mthbc[4];        /* 1 extrapolation, 2 periodic, 3 wall (bc1.f90, bc2.f90) */
    double xlower[2];
    double dx[2];
    int physics;         /* SC_PHYS_* */
    int efix;            /* -1: the reference solver's default; 0 off; 1 on */
    double gamma;        /* Euler (/param/ gamma; gamma1 = gamma-1) */
    double grav;         /* shallow water (/comrp/ grav) */
    double u;            /* advection speed (/comrp/ u) */
    double rho, bulk;    /* acoustics (/cparam/: cc = sqrt(bulk/rho), zz = cc*rho) */
    int precision;       /* SC_FP64 or SC_FP32 */
    int device;          /* HIP device ordinal */
    int mthlim[4];       /* limiter per wave family (sharpclaw.data mthlim; per
                            characteristic field 1..meqn for tvd2_char), read
                            when lim_type = 1: 1 minmod, 2 superbee, 3 van Leer,
                            4 MC, 5 Cada-Torrilhon (reconstruct.f90:544-571) */
    double tperiod;      /* 2D advection (velocities in aux(:,:,1:2)): before each
                            new time step the velocities are aux0 * cos(2 pi
                            (t + dt/2) / tperiod), aux0 the sc_set_aux array
                            (apps/advection/2d/swirl/b4step.f90:26-41); 0: constant */
} sc_params;

/* Time-step control of src/sharpclaw.f90:2-243 (read from sharpclaw.data). */
typedef struct sc_run_ctl {
    double dt_max;
    double cfl_max;
    double cfl_desired;
    int max_steps;
    int dt_variable;     /* 1: variable dt.  qold is refreshed every step in
                            both modes (the reference only does so for 1). */
} sc_run_ctl;

/* What sharpclaw() returns to main (sharpclaw.f90:230-235) plus counters. */
typedef struct sc_run_stats {
    double t;            /* time reached (tend on success) */
    double cfl_maxused, cfl_last;
    double dt_minused, dt_maxused, dt_last;
    int steps_taken;
    int info;            /* 0, 3 (cfl_desired > cfl_max), 11 (max_steps) */
    int rejected;        /* rejected step attempts */
    long long stage_evals; /* semi-discrete evaluations, rejected ones included */
    double gpu_ms;       /* device time of the stage work (HIP events) */
} sc_run_stats;

typedef struct sc_solver sc_solver;

const char *sc_last_error(void);
int sc_version(void);

/* ---- solver handle: device-resident q, qold and RK registers ---------- */

/* Allocates q, qold and the RK registers on params->device
 * (replaces qalloc, src/1d/qalloc1.f90:10-29, src/2d/qalloc2.f90:11-30). */
int sc_create(const sc_params *params, sc_solver **out);
int sc_destroy(sc_solver *s);

/* Host <-> device copies of the whole q array, ghosts included. */
int sc_set_q(sc_solver *s, const double *q);
int sc_get_q(sc_solver *s, double *q);

/* The aux array aux(1-mbc:mx+mbc, maux) of a 1D grid, or
 * aux(1-mbc:nx+mbc, 1-mbc:ny+mbc, maux) of a 2D grid, column-major as
 * setaux(maxmx,mbc,nx,xlower,dx,maux,aux) fills it (main.f90:164-166).
 * The device keeps the capacity field mcapa (flux1.f90:56-61; in 2D each
 * sweep's slice of it, flux2.f90:37-40,70-73), ghost cells included: the
 * ghost rows and columns count in the CFL with their own capacities; or 2D
 * advection's edge velocities aux(:,:,1:2).  The device Riemann solvers of
 * the other physics take no aux.  Needed before stepping when mcapa > 0.
 * 2D capacity runs take weno5 / weno5_char, mthbc 1-3 and tile 0 (16x15). */
int sc_set_aux(sc_solver *s, const double *aux);

/* Frame output, replacing output(meqn,mbc,nx,xlower,dx,q,t,iframe,aux,maux)
 * (main.f90:172,233; src/1d/out1.f90:32-104, src/2d/out2.f90:26-111) on the
 * data side: sc_output_begin sets interior values with |q| < 1e-99 to zero
 * in the solver's q itself, as the reference's writers do before printing
 * (out1.f90:63-68, out2.f90:60-67), fills the ghost cells, snapshots the
 * grid on the device and starts its copy to pinned host memory on a second
 * stream; it returns without waiting, and the solver may be advanced
 * meanwhile.  sc_output_wait blocks until that copy is done and gives the
 * grid in the sc_get_q layout.  One frame may be in flight per solver
 * (SC_ERR_STATE otherwise).  Writing fort.q/fort.t stays with the caller. */
int sc_output_begin(sc_solver *s);
int sc_output_wait(sc_solver *s, double *q);

/* Binary restart file: a header (grid shape, physics, t, the next dt) and
 * q in the sc_get_q layout with an FNV-1a checksum, written to path.tmp and
 * renamed.  sc_read_restart checks the header against the solver and the
 * checksum, uploads q and returns t and the next dt; continuing from it
 * reproduces the uninterrupted run bit for bit.  No counterpart in the
 * reference (it has no restart); for long runs that must survive a
 * process ending. */
int sc_write_restart(sc_solver *s, const char *path, double t, double dt_next);
int sc_read_restart(sc_solver *s, const char *path, double *t, double *dt_next);

/* qold = q (sharpclaw.f90:134) and q = qold (sharpclaw.f90:201).  Both are
 * buffer renamings on the device; no data moves. */
int sc_save_qold(sc_solver *s);
int sc_restore_qold(sc_solver *s);

/* step(q,g,dq,aux,dt,cfl,t,rp,src,tfluct) (src/1d/step1d.f90:2-31,
 * src/2d/step2d.f90:2-41): fills q's ghost cells in place (bc), sets
 * dq = dt*L(q) in the interior and returns the CFL number.  dq may be NULL
 * (result stays on the device); otherwise it receives the full array with
 * zero ghost entries. */
int sc_step(sc_solver *s, double dt, double t, double *dq, double *cfl);

/* flux1(q,g,dq,aux,dt,cfl,t,rp,tfluct,1) (src/flux1.f90:2-190) on a 1D grid
 * and flux2(q,g,dq,aux,dt,cfl,t,rp,tfluct) (src/2d/flux2.f90:2-94) on a 2D
 * grid, as step1d.f90:26-29 / step2d.f90:31-36 call them after the user's
 * bc and src: q is the caller's whole array (sc_get_q layout) with its
 * ghost cells as the caller's bc filled them, dq holds the caller's src
 * result and receives dq(interior) += dt*L(q); cfl receives the CFL number,
 * in 2D over the ghost rows and columns too (flux2.f90:36,68).  The solver's
 * q and qold are not touched.  This is the step of grids with a user
 * boundary (mthbc 0), for which sc_step, sc_ts and sc_advance return
 * SC_ERR_UNSUPPORTED, and of any caller that keeps q on the host. */
int sc_flux(sc_solver *s, const double *q, double *dq, double dt, double t, double *cfl);

/* sc_step on arrays the caller already holds on the device, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  q: device array in the sc_get_q
 * layout, its ghost cells filled in place (bc); dq: device array of the same
 * shape receiving dt*L(q) with zero ghost entries; cfl: one device double
 * receiving the CFL number; stream: a hipStream_t (NULL: the null stream).
 * The call only enqueues work on that stream (no host synchronisation; the
 * results are there when the stream reaches them).  Its only memory beyond
 * the arguments is one CFL word set (8 KiB) in a static device array per
 * device, kept until the process ends and cleared by each call: calls that
 * overlap on one device race.  Grids of any size; options that need a
 * grid-sized buffer or aux (2D slice-sweep reconstructions: lim_type 1, or
 * char_decomp 1 or 3; capacity; 2D advection) return SC_ERR_UNSUPPORTED. */
int sc_step_noalloc(const sc_params *p, double *q, double *dq, double dt, double t, double *cfl,
                    void *stream);

/* One SSP Runge-Kutta time step from q (= qold after sc_save_qold) with
 * every stage fused: ghost fill + flux + stage combination in one kernel
 * (ts_ssp22/42/33/104, src/sharpclaw.f90:247-464).  cfl receives the last
 * stage's CFL number, as the reference's loop uses it (sharpclaw.f90:184-216). */
int sc_ts(sc_solver *s, double dt, double told, double *cfl);

/* The whole time loop of sharpclaw(tstart,tend,...) (sharpclaw.f90:2-243):
 * dt clamp to tend, qold save, SSP step, CFL-driven dt update and
 * reject/retake.  *dt is dt_initial on entry and dt_last on return, as
 * main.f90 carries it from frame to frame. tend < tstart means one step. */
int sc_advance(sc_solver *s, const sc_run_ctl *ctl, double tstart, double tend,
               double *dt, sc_run_stats *stats);

/* Conservation sums  sum_cells q(.,m) * dx [* dy]  per component, in float64. */
int sc_sums(sc_solver *s, double *sums);

/* Device time of the fused stage launches of n back-to-back time steps of
 * params->time_integrator from the current q (every step restarts from q;
 * no CFL readback), in ms per stage launch, measured with HIP events on the
 * solver's stream.  Leaves q unchanged. */
int sc_time_stage(sc_solver *s, int n, double dt, double *ms_per_stage);

/* Per-launch device time of the fused stage kernels (no counterpart in the
 * reference; measurement).  sc_stage_timing(s, n) with n > 0: the next n
 * stage launches of sc_ts and sc_advance record their kernels' device start
 * and end in HIP events (hipExtLaunchKernelGGL; no host work per launch;
 * sc_advance takes its host-driven loop while timing is on); n = 0 turns it
 * off.  sc_stage_timing_get waits for the recorded launches and returns the
 * sum of their durations in ms and their number. */
int sc_stage_timing(sc_solver *s, int max_launches);
int sc_stage_timing_get(sc_solver *s, double *ms, long long *launches);

/* Tile shape of the 2D fused stage kernel (no counterpart in the reference):
 * -1 (default) times the candidates on this solver's grid at the first
 * sc_ts/sc_advance/sc_time_stage and keeps the fastest; 0..3 fix it
 * (16x16, 16x15, 15x16, 32x7 cells).  Every shape gives identical results.
 * The environment variable SC_TILE=k fixes it for new solvers. */
int sc_set_tile(sc_solver *s, int tile);
/* Current tile index (-1 before tuning) and its cell extents (tx, ty may be NULL). */
int sc_get_tile(sc_solver *s, int *tile, int *tx, int *ty);

/* Achievable HBM bandwidth: reps device-to-device copies of nbytes by a plain
 * copy kernel on `device`, timed with HIP events; *gbps = 2*nbytes*reps/time
 * (bytes read + written).  Calibrates the roofline and the HBM counters. */
int sc_copy_bench(int device, size_t nbytes, int reps, double *gbps);

/* ---- single routines on host arrays (each copies in, runs, copies out) */

/* rp(ixy,maxm,meqn,mwaves,mbc,mx,ql,qr,auxl,auxr,wave,s,amdq,apdq)
 * (e.g. apps/euler/shockentropy/rp1eu.f:5-6): i-th problem has left state
 * qr(i-1,:), right state ql(i,:), i = 2-mbc..mx+mbc.  Arrays are
 * (1-mbc:mx+mbc, ...) with mbc = params->mbc. */
int sc_rp(const sc_params *params, int ixy, int mx, const double *ql, const double *qr,
          double *wave, double *s, double *amdq, double *apdq);

/* evec(maxm,meqn,mbc,mx,q,auxl,auxr,evl,evr) as called by src/flux1.f90:105,
 * plus the sweep direction ixy: left/right eigenvectors at interface i-1/2,
 * i = 2-mbc..mx+mbc; evl(i,p,m), evr(i,m,p) are (1-mbc:mx+mbc, meqn, meqn). */
int sc_evec(const sc_params *params, int ixy, int mx, const double *q, double *evl,
            double *evr);

/* weno5(q,ql,qr) / weno5_char(q,ql,qr,evl,evr) (src/reconstruct.f90:77-143,
 * 146-241) on one slice of mx2 = mx+2*mbc cells (1-based, assumed shape).
 * Writes ql(i), qr(i-1) for the interfaces i = mbc+1..mx2-2 (entries outside
 * that range are left unchanged). */
int sc_weno5(const sc_params *params, int mx2, const double *q, double *ql, double *qr);
int sc_weno5_char(const sc_params *params, int mx2, const double *q, const double *evl,
                  const double *evr, double *ql, double *qr);

/* flux1(q1d,g,dq1d,aux,dt,cfl,t,rp,tfluct,ixy) (src/flux1.f90:2-190) on one
 * slice q1d(1-mbc:mx+mbc, meqn): dq1d(1:mx) = -dt/dx(ixy) * (A-dq(i+1) +
 * A+dq(i) + in-cell fluctuations) and the slice CFL number.  dq1d has the
 * same shape as q1d; its ghost entries are set to zero. */
int sc_flux1(const sc_params *params, int ixy, int mx, const double *q1d, double *dq1d,
             double dt, double *cfl);

/* ---- batched solves: many independent 1D grids in one persistent launch --
 * No counterpart in the reference, which solves one grid per process; each
 * solve follows the reference's time loop (sharpclaw.f90:35-235) exactly as
 * sc_advance does, with all its control decisions taken on the device. */

typedef struct sc_batch sc_batch;

/* Per-solve outcome of sc_batch_run (the fields of sc_run_stats plus timing). */
typedef struct sc_batch_result {
    double t;
    double cfl_maxused, cfl_last;
    double dt_minused, dt_maxused, dt_last;
    int steps_taken, rejected, info;
    int buffer;               /* device buffer holding the final q (internal) */
    long long stage_evals;
    long long t_begin, t_end; /* device wall-clock ticks */
    double done_ms;           /* completion time after the launch began, ms */
} sc_batch_result;

/* nsolves 1D grids sharing params (physics, reconstruction, precision,
 * integrator, boundary options); solve i has mx[i] cells on
 * [xlower[i], xupper[i]].  Every reconstruction of flux1.f90:65-117 that
 * single grids take (weno5, weno5_char, weno5_wave, weno5_trans, tvd2_wave,
 * tvd2_char); no aux arrays or source terms. */
int sc_batch_create(const sc_params *params, int nsolves, const int *mx, const double *xlower,
                    const double *xupper, sc_batch **out);
int sc_batch_destroy(sc_batch *b);
/* q(1-mbc:mx+mbc, meqn) of solve i */
int sc_batch_set_q(sc_batch *b, int i, const double *q);
int sc_batch_get_q(sc_batch *b, int i, double *q);
/* Advance every solve from tstart to tend (dt_initial each) in one launch;
 * results[i] for solve i; gpu_ms: HIP-event time of the launch.  The first
 * run of a batch also times workgroup sizes 64..1024 (64..256 for the
 * reconstructions with neighbour data: weno5_wave, weno5_trans, tvd2_*) on
 * one step of every solve (states untouched) and keeps the one with the
 * shortest predicted makespan. */
int sc_batch_run(sc_batch *b, const sc_run_ctl *ctl, double tstart, double tend, double dt_initial,
                 sc_batch_result *results, double *gpu_ms);
/* workgroup size chosen by the tuner (0 before the first run) */
int sc_batch_block(sc_batch *b, int *block);
/* the tuner's candidates and predicted whole-run times (ms; -1 for a size
 * not built for the batch's options); *n <= max */
int sc_batch_tuning(sc_batch *b, int max, int *blocks, double *predicted_ms, int *n);
/* Workgroups each solve ran on in the last run (parts[i] for solve i; 1: one
 * workgroup).  The tuner splits the largest solves over several co-resident
 * workgroups, which meet at a device barrier after every stage, when that
 * shortens the predicted makespan; the results are identical either way.
 * The environment variable SC_BATCH_PARTS=k caps the split (1: none). */
int sc_batch_parts(sc_batch *b, int *parts);

/* ---- request server: continuous batching on one device -----------------
 * No counterpart in the reference (one solve per process there); each
 * request runs the reference's time loop (sharpclaw.f90:35-235) on its own
 * grid, exactly as sc_advance / sc_batch_run do.  Requests are admitted in
 * arrival order while their device footprint fits the server's memory
 * budget and an active slot is free.  Admitted requests advance together in
 * rounds: per round, 1D requests with equal options share one launch of the
 * persistent batch kernel (a bounded number of time-loop iterations each),
 * and 2D requests with equal options advance K steps with every stage
 * launch covering all their grids' tiles and the time-loop decisions taken
 * on the device; loop state carries across rounds.  Arrivals join at the
 * next round.  A request's result does not depend on what else is in
 * flight.  Boundary options 1-3 only (ghost cells are synthesised). */

typedef struct sc_server sc_server;

typedef struct sc_request_stats {
    sc_run_stats run;        /* the time loop's record (gpu_ms unused) */
    double submit_ms;        /* server clock: ms since sc_server_create */
    double admit_ms;
    double done_ms;
    long long footprint;     /* device bytes held while active */
    int rounds;              /* rounds the request took part in */
} sc_request_stats;

typedef struct sc_server_stats {
    long long submitted, completed, rounds;
    long long peak_bytes, capacity_bytes, used_bytes;
    int active, queued, peak_active;
    double clock_ms;
} sc_server_stats;

/* hbm_bytes: device memory the server holds for request state (one
 * allocation); max_active: most requests in flight at once. */
int sc_server_create(int device, size_t hbm_bytes, int max_active, sc_server **out);
/* Queue a solve of q (copied) from tstart to tend; *id names it. */
int sc_server_submit(sc_server *s, const sc_params *params, const sc_run_ctl *ctl, double tstart,
                     double tend, double dt_initial, const double *q, long long *id);
/* Ids of requests finished since the last poll (at most max_ids). */
int sc_server_poll(sc_server *s, long long *ids, int max_ids, int *n);
/* Block until a finished request awaits polling or timeout_ms passes. */
int sc_server_wait(sc_server *s, double timeout_ms, int *n_ready);
/* Final q (with ghost cells filled) and record of a finished request, which
 * is then released. */
int sc_server_result(sc_server *s, long long id, double *q, sc_request_stats *st);
/* Block until every submitted request has finished (or timeout_ms). */
int sc_server_drain(sc_server *s, double timeout_ms);
int sc_server_get_stats(sc_server *s, sc_server_stats *st);
/* Unfinished requests are abandoned. */
int sc_server_destroy(sc_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SHARPCLAW_AMD_H */
