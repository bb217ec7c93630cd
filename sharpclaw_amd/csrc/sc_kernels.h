// Kernel arguments and host-side launchers of sc_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>

#include "sc_physics.h"

namespace sc {

// How a fused stage combines q^n (qold), the stage input q, r = dt*L(q)
// and the SSP(10,4) register qb (src/sharpclaw.f90:247-464).
enum StageOp {
    OP_DQ = 0,        // out = r                                (step: dq = dt*L(q))
    OP_ADD1 = 1,      // out = q + r                            (SSP22/33 stage 1)
    OP_SSP33_2 = 2,   // out = 0.75*qold + 0.25*(q + r)         (sharpclaw.f90:365)
    OP_SSP33_3 = 3,   // out = (qold + 2*(q + r))/3             (sharpclaw.f90:374)
    OP_SSP22_2 = 4,   // out = 0.5*(qold + q + r)               (sharpclaw.f90:274)
    OP_ADD_THIRD = 5, // out = q + r/3                          (sharpclaw.f90:304,313,322)
    OP_SSP42_4 = 6,   // out = 0.25*qold + 0.75*q + 0.25*r      (sharpclaw.f90:331)
    OP_ADD_SIXTH = 7, // out = q + r/6                          (sharpclaw.f90:396-455)
    OP_SSP104_5 = 8,  // r2 = q + r/6; qb = 0.04 qold + 0.36 r2; out = 0.6 qold + 0.4 r2 (424-426)
    OP_SSP104_10 = 9, // out = qb + 0.6*q + 0.1*r               (sharpclaw.f90:462)
};

enum TileOrder { TILE_BANDS = 0, TILE_ADAPTIVE = 1 };

struct StageArgs {
    const double *qin;   // stage input, grid layout with ghosts
    const double *qold;  // q^n
    const double *qbin;  // SSP(10,4) register read by OP_SSP104_10
    double *qout;        // stage output (dq for OP_DQ)
    double *qbout;       // SSP(10,4) register written by OP_SSP104_5
    unsigned long long *smax;  // [0] max|s| x-sweeps, [1] y-sweeps; NULL = no CFL
    int nx, ny;          // interior cells
    int LX, LY;          // extents with ghosts (nx+2mbc, ny+2mbc)
    int bc[4];           // 0: ghosts stored in qin; 1 extrap, 2 periodic, 3 wall
    double dtdx, dtdy;
    int op;
    int ntx, nty;        // tiles per direction (2D)
    // capacity aux(:,mcapa) (1D) or aux(:,:,mcapa) (2D) in the layout of one
    // field of q (NULL: none); 1D only: the source term (0 none, 1 radial
    // shallow water) with the stage's dt and the cell-centre coordinates
    // xlo + (i - 1/2) dxc it needs
    const double *capa;
    int src;
    double dt, xlo, dxc;
    // 2D advection: edge velocities aux(:,:,1:2) in the q layout (plane
    // LX*LY each) and the step's factor on them (b4step)
    const double *aux;
    double vscale;
    // 2D: 1 = every tile takes the full path (no constant-tile shortcut);
    // the tile-shape tuner times the full path
    int full;
    // 2D: order of the tiles over the workgroups (TILE_*, tile_of_workgroup).
    // TILE_ADAPTIVE takes the tile list of the last tile_partition (tlist,
    // *tnf); every 2D stage writes each tile's flag (1: not constant) to
    // tflag when it is not NULL.
    int order;
    unsigned char *tflag;
    const int *tlist;
    const int *tnf;
    // 2D, stages 2 and 3 of a step: the flags the step's first stage wrote
    // (stage2d_tile's deep-constant tiles); NULL: not used
    const unsigned char *dflag;
};

// physics selector of the launchers (sc_physics.h PhysId)
struct PhysSel {
    int phys;     // PhysId
    int chr;      // char_decomp
    int fp32;     // 0 float64, 1 float32 arithmetic
    int lim = 2;  // lim_type
};

// start / stop events the next stage launch (launch_stage1d, launch_stage2d,
// launch_sweep2d) of the calling host thread reports its device times to;
// nullptr, nullptr for plain launches
void set_launch_events(hipEvent_t start, hipEvent_t stop);
hipError_t launch_stage1d(const PhysSel &s, int ixy, const StageArgs &a, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st);
// 2D stage by x and y slice sweeps (the reconstructions the tile kernel does
// not stage: lim_type 1, char_decomp 1 or 3); dqx: q-sized scratch
bool sweep2d_path(const PhysSel &s);
hipError_t launch_sweep2d(const PhysSel &s, const StageArgs &a, double *dqx, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st);
// 2D tile shapes selectable per launch (index into a fixed table)
constexpr int NTILE2D = 4;
hipError_t launch_stage2d(const PhysSel &s, int tile, const StageArgs &a, const PhysConst<double> &cd,
                          const PhysConst<float> &cf, hipStream_t st);
void tile2d_shape(int tile, int *tx, int *ty);
// The tile list of TILE_ADAPTIVE from the flags of the last stage launch: the
// tiles that were not constant, then the constant ones, each in strip order
// (strip_tile); *nf = the number of the former.  launch_tile_list_reset: no
// flags, the list in strip order.  With set != NULL it also writes the CFL
// set's two words to out (as launch_cfl_to_host).  bcount: ceil(ntiles /
// 1024) ints of scratch (grids of more than 16384 tiles take two kernels).
hipError_t launch_tile_partition(const unsigned char *flag, int ntx, int nty, int *list, int *nf,
                                 const unsigned long long *set, double *out, int *bcount, hipStream_t st);
// the max over the copies of the x and y words of a CFL set into out[0..1]
hipError_t launch_cfl_to_host(const unsigned long long *set, double *out, hipStream_t st);
// The device-resident loop of one 2D grid (Res2D, sc_batch2d.h; grids of at
// most 16384 tiles): stage k of the current iteration on every tile (a: the
// stage's fixed arguments; its buffers and dt come from R), a no-op once
// the loop is done; and the control launch between iterations (do_post:
// close the iteration that ran -- CFL, dt, accept/reject -- and rebuild the
// tile list from its flags; do_pre: open the next).
struct Res2D;
hipError_t launch_stage2d_res(const PhysSel &s, int tile, const StageArgs &a, const Res2D *R, int k,
                              const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st);
hipError_t launch_res2d_ctl(Res2D *R, const unsigned char *tflag, int *tlist, int *tnf, unsigned long long *set,
                            int ntx, int nty, int do_post, int do_pre, hipStream_t st);
hipError_t launch_tile_list_reset(unsigned char *flag, int ntx, int nty, int *list, int *nf, hipStream_t st);
// ghost fill in place (bc1.f90 / bc2.f90 semantics, corners included)
hipError_t launch_bc(int ndim, int meqn, double *q, int nx, int ny, int LX, int LY, const int *bc,
                     hipStream_t st);
// single routines on device arrays in the reference layouts
hipError_t launch_rp(const PhysSel &s, int ixy, int mx, int mbc, const double *ql, const double *qr,
                     double *wave, double *sp, double *amdq, double *apdq,
                     const PhysConst<double> &cd, const PhysConst<float> &cf, hipStream_t st);
hipError_t launch_evec(const PhysSel &s, int ixy, int mx, int mbc, const double *q, double *evl,
                       double *evr, const PhysConst<double> &cd, const PhysConst<float> &cf,
                       hipStream_t st);
hipError_t launch_weno(int meqn, int chr, int fp32, int mx2, const double *q, const double *evl,
                       const double *evr, double *ql, double *qr, hipStream_t st);
// deterministic two-pass sums of the interior, per component
hipError_t launch_sums(int ndim, int meqn, const double *q, int nx, int ny, int LX, int LY,
                       double *partial, int nblk, double *out, hipStream_t st);
hipError_t launch_copy(double *dst, const double *src, size_t n, hipStream_t st);
// |q| < 1e-99 -> 0 over the interior, in place (the reference's frame writers)
hipError_t launch_flush_small(int ndim, int meqn, double *q, int nx, int ny, int LX, int LY,
                              hipStream_t st);

}  // namespace sc
