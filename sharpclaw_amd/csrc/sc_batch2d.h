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

// Host side (sc_batch2d.hip).
hipError_t launch_stage2d_batch(const PhysSel &s, int tile, const Batch2DArgs &A, int ntiles,
                                const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st);
// do_post: close the iteration whose stages just ran; do_pre: open the next.
hipError_t launch_ctl2d(const Batch2DArgs &A, int do_post, int do_pre, hipStream_t st);
__host__ __device__ int stages_of_integrator(int ti);

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
