// Batched 2D grids advanced on the device: the stage launches of one time
// step cover the tiles of every active grid, and a control kernel takes
// each grid's time-loop decisions (sharpclaw.f90:35-235: dt clamp to tend,
// qold save, CFL-driven dt update, accept or reject-and-retake) from its
// CFL words, so a round of many steps runs without the host.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch_kernel.h"

namespace sc {

struct Grid2D {
    long long off;      // first double of this grid's 4 arrays (3 q buffers, SSP(10,4) register)
    int nx, ny, LX, LY;
    int ntx, nty;       // tiles per direction
    int tile0;          // first launch-wide tile index of this grid
    int maxn;           // iteration limit of the loop
    double dx, dy, tstart, tend;
    sc_run_ctl ctl;
};

struct Batch2DArgs {
    const Grid2D *grids;        // ascending tile0
    SolveState *state;
    unsigned long long *smax;   // one CFL set (CFL_WORDS words, x and y sweeps) per grid
    int ngrids;
    double *base;
    int bc[4];
    int ti;                     // time integrator 1..4
    int stage;                  // stage of a stage launch
};

__host__ __device__ int stages_of_integrator(int ti);

// The time-loop decisions of one 2D grid (sharpclaw.f90:35-235), taken on
// the device between steps.
__device__ __forceinline__ void finish_loop2d(const Grid2D &G, SolveState &s)
{
    if (s.info == 0 && G.ctl.dt_variable == 1 && s.t < G.tend && s.steps == G.maxn) s.info = 11;
    if (s.info == 0 && G.ctl.dt_variable == 0 && s.cflmax > G.ctl.cfl_max) s.info = 12;
    s.done = 1;
    s.go = 0;
    s.t_end = (long long)wall_clock64();
}

// Close the iteration whose nst stages just ran, from the max |s| of its
// last stage in x (sx) and y (sy): the CFL number (flux2.f90:47,80), the dt
// update and accept, or reject and retake (sharpclaw.f90:184-216).
__device__ __forceinline__ void loop2d_post(const Grid2D &G, SolveState &s, int nst, double sx, double sy)
{
    if (!s.go) return;
    s.go = 0;
    s.stages += nst;
    s.cfl = fmax((s.dt / G.dx) * sx, (s.dt / G.dy) * sy);
    const sc_run_ctl &ctl = G.ctl;
    if (ctl.dt_variable == 1) {
        if (s.cfl > 0.0) {
            s.dt = fmin(ctl.dt_max, s.dt * ctl.cfl_desired / s.cfl);
            s.dtmin = fmin(s.dt, s.dtmin);
            s.dtmax = fmax(s.dt, s.dtmax);
        } else {
            s.dt = ctl.dt_max;
        }
    }
    bool fin = false, counted = true;
    if (s.cfl <= ctl.cfl_max) {
        s.cflmax = fmax(s.cfl, s.cflmax);
        int in = s.cur, out = other_buf(in, s.old);
        for (int k = 1; k < nst; k++) {
            in = out;
            out = other_buf(in, s.old);
        }
        s.cur = out;
    } else {
        s.t = s.told;
        s.cur = s.old;  // q = qold
        s.rejected++;
        if (ctl.dt_variable == 1) {
            s.retake = 1;
        } else {
            s.cflmax = fmax(s.cfl, s.cflmax);
            fin = true;
            counted = false;
        }
    }
    if (counted) {
        s.steps++;
        if (s.t >= G.tend) fin = true;
    }
    if (!fin && (s.info != 0 || s.it >= G.maxn)) fin = true;
    if (fin) finish_loop2d(G, s);
}

// Open the next iteration unless the loop is done: the dt clamp to tend and
// qold = q at a new step (sharpclaw.f90:118-150).  Returns whether an
// iteration was opened (its CFL words are then due for clearing).
__device__ __forceinline__ bool loop2d_pre(const Grid2D &G, SolveState &s)
{
    if (s.done) return false;
    if (s.info != 0 || s.it >= G.maxn) {
        finish_loop2d(G, s);
        return false;
    }
    if (s.t_begin == 0) s.t_begin = (long long)wall_clock64();
    s.it++;
    if (!s.retake) {
        s.told = s.t;
        if (s.told + s.dt > G.tend && G.tstart < G.tend) s.dt = G.tend - s.told;
        s.old = s.cur;  // qold = q
    }
    s.retake = 0;
    s.t = s.told + s.dt;
    s.go = 1;
    return true;
}

// The device-resident time loop of one 2D grid (sc_advance): a control
// block in device memory from which the stage launches of a round take the
// iteration's dt, buffers and tile count (ResStage), rewritten by the
// control kernel between iterations, so that the host enqueues a round of
// iterations without waiting for any of them.
// What a stage launch of the iteration needs beyond its fixed arguments,
// in one 64-byte record (a single scalar load per workgroup).
struct __align__(64) ResStage {
    double dtdx, dtdy, dt;
    int in, old, out;         // buffers (0..2) of the stage's input, q^n and output
    int go;                   // 0: the loop is done, the launch does nothing
    int nf;                   // non-constant tiles at the head of the tile list
};

struct Res2D {
    ResStage st[10];          // the stages of the current iteration
    Grid2D G;                 // tstart, tend, maxn, dx, dy, ctl
    SolveState s;             // the loop state
    int nst;                  // stages per iteration
};

// Host side (sc_batch2d.hip).
hipError_t launch_stage2d_batch(const PhysSel &s, int tile, const Batch2DArgs &A, int ntiles,
                                const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st);
// do_post: close the iteration whose stages just ran; do_pre: open the next.
hipError_t launch_ctl2d(const Batch2DArgs &A, int do_post, int do_pre, hipStream_t st);

}  // namespace sc

namespace sc {

// Host driver of rounds over a set of 2D grids with equal options: it owns
// the device descriptor/state/CFL arrays and their pinned staging, and
// enqueues a round of K iterations (K x (stages + 1) + 1 launches) on its
// stream without waiting.
struct Batch2DRunner {
    PhysSel sel;
    PhysConst<double> cd;
    PhysConst<float> cf;
    int tile = 0;       // tile shape (sc_kernels.h tile table)
    int ti = 3;
    int bc[4] = {1, 1, 1, 1};
    hipStream_t st = nullptr;
    bool owns_st = true;
    hipEvent_t ev = nullptr;
    int cap = 0;
    Grid2D *d_grids = nullptr;
    SolveState *d_state = nullptr;
    unsigned long long *d_smax = nullptr;
    Grid2D *h_grids = nullptr;      // pinned
    SolveState *h_state = nullptr;  // pinned
    // a captured round (continue_round), valid for (g_n, g_K, g_tiles, g_base)
    hipGraphExec_t g_exec = nullptr;
    int g_n = 0, g_K = 0;
    long long g_tiles = 0;
    double *g_base = nullptr;

    // stream: the caller's stream to work on, or NULL for one of its own
    int init(const sc_params &p, hipStream_t stream = nullptr);
    void set_tile(int t);
    int reserve(int n);
    // h_grids[0..n) and h_state[0..n) filled by the caller (tile0 is set
    // here): uploaded, then one round; states come back into h_state when
    // the round's event fires.
    int launch_round(int n, int K, double *base);
    // One more round from the device states left by the previous round,
    // replayed from a captured graph when the caller allows it.
    int continue_round(int n, int K, double *base, bool graph);
    int wait();
    void destroy();

  private:
    long long assign_tiles(int n);
    hipError_t enqueue_steps(int n, int K, long long tiles, double *base);
};
// Iterations per round so that a round is a few milliseconds of work.
int round_iterations(int stages, long long tiles);

}  // namespace sc
