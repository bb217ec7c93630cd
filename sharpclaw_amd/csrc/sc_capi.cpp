// C ABI of sharpclaw_amd (include/sharpclaw_amd.h).
//
// The solver handle owns the device copies of q, qold and the RK registers
// that module Global / type rkreg hold in the reference
// (src/1d/Global1.f90, src/1d/qalloc1.f90:10-29).  Three q-sized buffers
// rotate through the stages so that "qold = q" and "q = qold"
// (sharpclaw.f90:134,201) are renamings; a fourth holds the SSP(10,4)
// register r1 (sharpclaw.f90:425), a fifth dq for sc_step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>
#include <cstdint>

#include "../../include/sharpclaw_amd.h"
#include "sc_batch2d.h"
#include "sc_batch_host.h"
#include "sc_internal.h"
#include "sc_kernels.h"

using namespace sc;

static thread_local std::string g_err;

int sc_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return sc_fail(SC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));          \
    } while (0)

extern "C" const char *sc_last_error(void) { return g_err.c_str(); }
extern "C" int sc_version(void) { return 1; }

// --------------------------------------------------------- validation

int sc_phys_id(const sc_params *p)
{
    switch (p->physics) {
    case SC_PHYS_ADVECTION: return p->ndim == 1 ? ADV1 : ADV2;
    case SC_PHYS_ACOUSTICS: return p->ndim == 1 ? AC1 : -1;
    case SC_PHYS_EULER: return p->ndim == 1 ? EU1 : EU2;
    case SC_PHYS_SHALLOW_WATER: return p->ndim == 1 ? SW1 : SW2;
    }
    return -1;
}

static int shape_m(int id)
{
    static const int M[] = {1, 2, 3, 2, 4, 3, 1};
    return M[id];
}
static int shape_w(int id)
{
    static const int W[] = {1, 2, 3, 2, 3, 3, 1};
    return W[id];
}

// Options the device path implements; everything else is rejected with the
// reason, as the reference stops on an option it does not know.
int sc_check_params(const sc_params *p, bool grid)
{
    if (!p) return sc_fail(SC_ERR_ARG, "params is NULL");
    if (p->ndim != 1 && p->ndim != 2) return sc_fail(SC_ERR_ARG, "ndim must be 1 or 2 (got %d)", p->ndim);
    int id = sc_phys_id(p);
    if (id < 0) return sc_fail(SC_ERR_UNSUPPORTED, "physics %d has no %dD device solver", p->physics, p->ndim);
    if (p->meqn != shape_m(id) || p->mwaves != shape_w(id))
        return sc_fail(SC_ERR_ARG, "physics %d in %dD needs meqn=%d mwaves=%d (got %d, %d)", p->physics,
                    p->ndim, shape_m(id), shape_w(id), p->meqn, p->mwaves);
    if (p->mbc != 3) return sc_fail(SC_ERR_ARG, "mbc must be 3 for WENO5 (reconstruct.f90:85)");
    // aux arrays (main.f90:96-98): the device physics take no aux in their
    // Riemann solvers; a capacity column (flux1.f90:56-61) scales the update
    // and the CFL of 1D grids and of 2D Euler / shallow water (flux2.f90
    // passes each sweep's aux slice to flux1)
    if (p->maux < 0 || p->mcapa < 0 || p->mcapa > p->maux)
        return sc_fail(SC_ERR_ARG, "mcapa (%d) must index aux(:,1..maux) (maux = %d)", p->mcapa, p->maux);
    if (id == ADV2) {
        // apps/advection/2d/swirl: velocities in aux(:,:,1:2), char_decomp 0
        if (p->maux < 2 || p->mcapa != 0)
            return sc_fail(SC_ERR_ARG, "2D advection takes its edge velocities from aux(:,:,1:2): maux >= 2, mcapa = 0");
        if (p->char_decomp != 0 || p->lim_type != 2 || p->tfluct_solver != 0)
            return sc_fail(SC_ERR_UNSUPPORTED, "2D advection: weno5 (char_decomp 0, lim_type 2, tfluct_solver 0)");
    } else if (p->maux > 0 && p->ndim != 1) {
        if (p->mcapa == 0)
            return sc_fail(SC_ERR_UNSUPPORTED, "aux arrays in 2D: 2D advection's velocities, or a capacity (mcapa > 0)");
        if (p->lim_type != 2 || (p->char_decomp != 0 && p->char_decomp != 2))
            return sc_fail(SC_ERR_UNSUPPORTED, "a capacity in 2D runs with weno5 or weno5_char (lim_type 2, "
                                               "char_decomp 0 or 2)");
    }
    // reconstructions (flux1.f90:65-117)
    if (p->lim_type == 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "lim_type 0: q2qlqr_poly leaves the interfaces within k of the boundary "
                                           "unset (polyrecon.f90 loops i = k+1..mx-k-1)");
    if (p->lim_type != 1 && p->lim_type != 2) return sc_fail(SC_ERR_ARG, "lim_type must be 0, 1 or 2 (got %d)", p->lim_type);
    if (p->lim_type == 2 && (p->char_decomp < 0 || p->char_decomp > 3))
        return sc_fail(SC_ERR_ARG, "char_decomp must be 0..3 with lim_type 2 (flux1.f90:111-114)");
    if (p->lim_type == 1 && p->char_decomp == 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "lim_type 1 char_decomp 0: tvd2 reads its first slope before setting it "
                                           "and divides by zero jumps unguarded (reconstruct.f90:540-542)");
    if (p->lim_type == 1 && p->char_decomp != 1 && p->char_decomp != 2)
        return sc_fail(SC_ERR_ARG, "char_decomp must be 1 or 2 with lim_type 1 (flux1.f90:76-91)");
    // in 2D these run as x and y slice sweeps (flux2.f90), for Euler and
    // shallow water
    if ((p->lim_type == 1 || p->char_decomp == 1 || p->char_decomp == 3) && p->ndim == 2 && id != EU2 && id != SW2)
        return sc_fail(SC_ERR_UNSUPPORTED, "lim_type %d char_decomp %d: implemented for 1D grids and 2D Euler / "
                       "shallow water", p->lim_type, p->char_decomp);
    // tvd2_wave limits wave families 1..mwaves, tvd2_char characteristic
    // fields 1..meqn (reconstruct.f90:623, 726; with meqn > mwaves, 2D Euler,
    // the reference reads mthlim past its mwaves entries)
    if (p->lim_type == 1)
        for (int k = 0; k < (p->char_decomp == 2 ? p->meqn : p->mwaves) && k < 4; k++)
            if (p->mthlim[k] < 1 || p->mthlim[k] > 5)
                return sc_fail(SC_ERR_ARG, "mthlim(%d) = %d: limiters are 1..5 (reconstruct.f90:626-654)", k + 1,
                               p->mthlim[k]);
    if (p->tfluct_solver != 0 && p->tfluct_solver != 1)
        return sc_fail(SC_ERR_ARG, "tfluct_solver must be 0 or 1 (got %d)", p->tfluct_solver);
    // tfluct routines exist for advection, acoustics and 1D shallow water
    // (apps/advection/1d/tfluct1.f, apps/acoustics/1d/homogeneous/tfluct.f,
    // apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/tfluct1sw.f); the
    // library's own tfluct (src/tfluct.f90) stops the run
    if (p->tfluct_solver == 1 && (p->ndim != 1 || p->physics == SC_PHYS_EULER))
        return sc_fail(SC_ERR_UNSUPPORTED, "tfluct_solver 1: the reference has a tfluct routine for 1D advection, acoustics and shallow water only");
    // source terms: 0 none (src/1d/src1.f90, src/2d/src2.f90); 1 the radial
    // shallow-water source of apps/shallowWater/2dRadialDam/1dRadialSymm-qwave/src1.f90
    if (p->src_term == 1 && !(p->physics == SC_PHYS_SHALLOW_WATER && p->ndim == 1))
        return sc_fail(SC_ERR_UNSUPPORTED, "src_term 1 (radial shallow-water source) is for 1D shallow water");
    if (p->src_term != 0 && p->src_term != 1)
        return sc_fail(SC_ERR_UNSUPPORTED, "src_term %d: device sources are 0 (none) and 1 (radial shallow water)",
                       p->src_term);
    if (p->precision != SC_FP64 && p->precision != SC_FP32) return sc_fail(SC_ERR_ARG, "precision must be SC_FP64 or SC_FP32");
    if (!grid) return SC_OK;
    if (p->time_integrator < 1 || p->time_integrator > 4)
        return sc_fail(SC_ERR_ARG, "time_integrator should be between 1 and 4 (sharpclaw.f90:172-174)");
    for (int d = 0; d < p->ndim; d++) {
        if (p->nx[d] < 1) return sc_fail(SC_ERR_ARG, "nx[%d] = %d", d, p->nx[d]);
        if (!(p->dx[d] > 0.0)) return sc_fail(SC_ERR_ARG, "dx[%d] must be positive", d);
        // mthbc 0: the user's bc fills the ghost cells (bc1.f90:34-37,
        // bc2.f90:35,125); such grids step through sc_flux, which takes them
        // from the caller's array
        int bl = p->mthbc[2 * d], br = p->mthbc[2 * d + 1];
        if (bl < 0 || bl > 3 || br < 0 || br > 3) return sc_fail(SC_ERR_ARG, "mthbc must be 0, 1, 2 or 3");
        if ((bl == 2) != (br == 2))
            return sc_fail(SC_ERR_ARG, "periodic boundary conditions require mthbc(%d) and mthbc(%d) both be 2 (main.f90:140-145)",
                        2 * d + 1, 2 * d + 2);
        if ((bl == 2 || bl == 3 || br == 3) && p->nx[d] < 3)
            return sc_fail(SC_ERR_ARG, "periodic/wall boundaries need at least 3 cells");
        int mnorm = d == 0 ? 2 : 3;
        if ((bl == 3 || br == 3) && p->meqn < mnorm)
            return sc_fail(SC_ERR_ARG, "a wall negates component %d, which meqn=%d does not have", mnorm, p->meqn);
    }
    return SC_OK;
}

// ------------------------------------------------------------- handle

static constexpr int NSLOT = 64;  // CFL sets (CFL_WORDS words each) before the ring is cleared

struct sc_solver {
    sc_params p;
    PhysSel sel;
    PhysConst<double> cd;
    PhysConst<float> cf;
    int dev;
    hipStream_t st;
    size_t n;          // doubles per grid array
    double *buf[3];    // rotating q buffers
    double *qb;        // SSP(10,4) register
    double *dq;        // step() output
    double *dqx;       // 2D slice sweeps: dq of the x sweeps (NULL on the tile path)
    double *partial;   // sums scratch
    double *sums;      // sums result
    int nblk_sums;
    unsigned long long *slots;  // CFL words (device)
    double *hcfl;      // pinned: the CFL set's x and y maxima (read_cfl)
    long long slot_ctr;
    int cur, old;      // buffer indices of q and qold
    bool have_q;
    int tile2d;        // 2D tile shape index; -1: tune at first use
    int tile_order;    // 2D tile order over the workgroups (StageArgs::order)
    // TILE_ADAPTIVE: per-tile flags of the last stage, the tile list built
    // from them after each step and its count of non-constant tiles, for the
    // tile shape tlist_tile (-1: not built)
    unsigned char *tflag;
    unsigned char *dflag;  // flags of a step's first stage (the deep-constant tiles of its later stages)
    int *tlist, *tnf, *tcount;
    int tlist_tile;
    double *mem;       // one allocation: buf[0..2], qb, dq[, dqx]
    int loop_mode;     // sc_advance: -1 by grid size, 0 host-driven, 1 device-resident
    bool use_graph;    // device loop: replay rounds from a captured graph
    Batch2DRunner *dr; // 2D device-resident loop replayed from a graph (SC_GRAPH; created at first use)
    // 2D device-resident loop (created at first use): the control block and
    // its CFL set on the device, a pinned copy for the upload and the state readback
    Res2D *res;
    unsigned long long *rset;
    Res2D *hres;
    hipEvent_t ev0, ev1;
    double *capa;      // capacity aux(:,mcapa) in the q layout (1D, mcapa > 0)
    double *aux2;      // 2D advection: edge velocities aux(:,:,1:2)
    double *gsl;       // sc_flux on 2D grids: the four ghost slices and their dq (created at first use)
    double auxg[2];    // their max |.| over the ghost rows (x) / columns (y) swept for the CFL
    double vscale;     // b4step's factor on the velocities for the current step
    bool vt_stale;     // a new step starts: vscale is due
    // frame output (created at first sc_output_begin): device snapshot,
    // pinned host copy, copy stream and its events
    double *snap;
    double *hsnap;
    hipStream_t cst;
    hipEvent_t ev_snap, ev_out;
    bool out_pending;
    // per-launch timing of the fused stage kernels (sc_stage_timing): the
    // start / stop event pairs of the timed launches, read at sc_stage_timing_get
    bool timing;
    hipEvent_t *tev;
    int tev_cap, tev_used;
    // 1D device-resident loop (created at first use): the grid as one solve
    // of the persistent batch kernel, split over its chunks
    struct Loop1D {
        BatchSolve *solve;
        SolveState *state;
        int *order;
        unsigned long long *next, *sync;
        int ncu;
        double tick_ms;
        double *priv;       // private stage buffers of the parts (whole steps)
        size_t priv_bytes;
    } l1;
};

struct DevGuard {
    int prev;
    explicit DevGuard(int d) { hipGetDevice(&prev); hipSetDevice(d); }
    ~DevGuard() { hipSetDevice(prev); }
};

static size_t grid_n(const sc_params *p)
{
    size_t n = (size_t)(p->nx[0] + 2 * p->mbc) * p->meqn;
    if (p->ndim == 2) n *= (size_t)(p->nx[1] + 2 * p->mbc);
    return n;
}

extern "C" int sc_create(const sc_params *params, sc_solver **out)
{
    if (!out) return sc_fail(SC_ERR_ARG, "out is NULL");
    *out = nullptr;
    int rc = sc_check_params(params, true);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (params->device < 0 || params->device >= ndev)
        return sc_fail(SC_ERR_ARG, "device %d out of range (%d devices)", params->device, ndev);
    DevGuard g(params->device);
    sc_solver *s = new sc_solver();
    memset((void *)s, 0, sizeof(*s));
    s->p = *params;
    if (s->p.ndim == 1) {
        s->p.nx[1] = 1;
        s->p.dx[1] = 1.0;
    }
    s->sel.phys = sc_phys_id(params);
    s->vscale = 1.0;  // the velocities as uploaded until the first step
    s->vt_stale = true;
    s->sel.chr = params->char_decomp;
    s->sel.lim = params->lim_type;
    s->sel.fp32 = params->precision == SC_FP32;
    s->cd = sc_make_const<double>(params);
    s->cf = sc_make_const<float>(params);
    s->dev = params->device;
    s->n = grid_n(&s->p);
    s->nblk_sums = 256;
    const bool sweeps = s->p.ndim == 2 && sweep2d_path(s->sel);
    const int narr = sweeps ? 6 : 5;
    bool ok = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipMalloc(&s->mem, narr * s->n * sizeof(double)) == hipSuccess;
    if (ok) {
        for (int k = 0; k < 3; k++) s->buf[k] = s->mem + k * s->n;
        s->qb = s->mem + 3 * s->n;
        s->dq = s->mem + 4 * s->n;
        s->dqx = sweeps ? s->mem + 5 * s->n : nullptr;
    }
    ok = ok && hipMalloc(&s->partial, (size_t)s->nblk_sums * s->p.meqn * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&s->sums, s->p.meqn * sizeof(double)) == hipSuccess;
    ok = ok && hipMalloc(&s->slots, NSLOT * CFL_WORDS * sizeof(unsigned long long)) == hipSuccess;
    ok = ok && hipHostMalloc(&s->hcfl, 2 * sizeof(double), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipEventCreate(&s->ev0) == hipSuccess && hipEventCreate(&s->ev1) == hipSuccess;
    if (ok) {
        for (int k = 0; k < 3 && ok; k++)
            ok = hipMemsetAsync(s->buf[k], 0, s->n * sizeof(double), s->st) == hipSuccess;
        ok = ok && hipMemsetAsync(s->qb, 0, s->n * sizeof(double), s->st) == hipSuccess;
        ok = ok && hipMemsetAsync(s->dq, 0, s->n * sizeof(double), s->st) == hipSuccess;
        if (s->dqx) ok = ok && hipMemsetAsync(s->dqx, 0, s->n * sizeof(double), s->st) == hipSuccess;
        ok = ok && hipMemsetAsync(s->slots, 0, NSLOT * CFL_WORDS * sizeof(unsigned long long), s->st) == hipSuccess;
        ok = ok && hipStreamSynchronize(s->st) == hipSuccess;
    }
    if (!ok) {
        sc_destroy(s);
        return sc_fail(SC_ERR_NOMEM, "device allocation of %zu doubles x %d failed", grid_n(params), narr);
    }
    s->cur = 0;
    s->old = 0;
    s->tile2d = -1;
    s->loop_mode = -1;
    if (const char *e = getenv("SC_HOST_LOOP")) s->loop_mode = atoi(e) != 0 ? 0 : 1;
    s->use_graph = false;
    if (const char *e = getenv("SC_GRAPH")) s->use_graph = atoi(e) != 0;
    s->tile_order = TILE_ADAPTIVE;
    if (const char *e = getenv("SC_TILE_ORDER")) s->tile_order = atoi(e) == 0 ? TILE_BANDS : TILE_ADAPTIVE;
    s->tlist_tile = -1;
    if (s->p.ndim == 2 && !sweeps) {
        // tile lists for the shape with the most tiles
        int nmax = 0;
        for (int k = 0; k < NTILE2D; k++) {
            int tx, ty;
            tile2d_shape(k, &tx, &ty);
            nmax = std::max(nmax, ((s->p.nx[0] + tx - 1) / tx) * ((s->p.nx[1] + ty - 1) / ty));
        }
        void *tm = nullptr;
        const int nbc = (nmax + 1023) / 1024;
        if (hipMalloc(&tm, (size_t)(nmax + nbc + 2) * sizeof(int) + 2 * (size_t)nmax) != hipSuccess) {
            sc_destroy(s);
            return sc_fail(SC_ERR_NOMEM, "device allocation of the tile lists (%d tiles) failed", nmax);
        }
        s->tnf = (int *)tm;
        s->tlist = s->tnf + 2;
        s->tcount = s->tlist + nmax;
        s->tflag = (unsigned char *)(s->tcount + nbc);
        s->dflag = s->tflag + nmax;
    }
    if (const char *e = getenv("SC_TILE")) {
        int k = atoi(e);
        // a 2D capacity runs on tile 0, the shape its kernel variant is built for
        if (k >= 0 && k < NTILE2D && !(s->p.ndim == 2 && s->p.mcapa > 0)) s->tile2d = k;
    }
    *out = s;
    return SC_OK;
}

// The frame-output resources (sc_output_begin), released and reset
// together, so that a partly failed creation is retried whole.
static void output_free(sc_solver *s)
{
    if (s->cst) hipStreamSynchronize(s->cst);
    if (s->snap) hipFree(s->snap);
    if (s->hsnap) hipHostFree(s->hsnap);
    if (s->ev_snap) hipEventDestroy(s->ev_snap);
    if (s->ev_out) hipEventDestroy(s->ev_out);
    if (s->cst) hipStreamDestroy(s->cst);
    s->snap = nullptr;
    s->hsnap = nullptr;
    s->ev_snap = nullptr;
    s->ev_out = nullptr;
    s->cst = nullptr;
}

extern "C" int sc_destroy(sc_solver *s)
{
    if (!s) return SC_OK;
    DevGuard g(s->dev);
    if (s->st) hipStreamSynchronize(s->st);
    if (s->dr) {
        s->dr->destroy();
        delete s->dr;
    }
    if (s->mem) hipFree(s->mem);
    if (s->partial) hipFree(s->partial);
    if (s->sums) hipFree(s->sums);
    if (s->slots) hipFree(s->slots);
    if (s->hcfl) hipHostFree(s->hcfl);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    if (s->capa) hipFree(s->capa);
    if (s->aux2) hipFree(s->aux2);
    if (s->gsl) hipFree(s->gsl);
    if (s->res) hipFree(s->res);
    if (s->rset) hipFree(s->rset);
    if (s->hres) hipHostFree(s->hres);
    if (s->tnf) hipFree(s->tnf);
    if (s->l1.solve) hipFree(s->l1.solve);
    if (s->l1.priv) hipFree(s->l1.priv);
    for (int k = 0; k < s->tev_cap; k++)
        if (s->tev[k]) hipEventDestroy(s->tev[k]);
    free(s->tev);
    output_free(s);
    if (s->st) hipStreamDestroy(s->st);
    delete s;
    return SC_OK;
}

// Steps of grids with a user boundary (mthbc 0) need the caller's bc before
// every stage: they go through sc_flux.
static int user_bc_refused()
{
    return sc_fail(SC_ERR_UNSUPPORTED, "mthbc 0 (user boundary conditions): the caller's bc fills the ghost "
                                       "cells before every stage; step with sc_flux (the flux1/flux2 drop-in)");
}

static int free_buf(const sc_solver *s, int a, int b)
{
    for (int k = 0; k < 3; k++)
        if (k != a && k != b) return k;
    return -1;
}

extern "C" int sc_set_q(sc_solver *s, const double *q)
{
    if (!s || !q) return sc_fail(SC_ERR_ARG, "NULL argument");
    DevGuard g(s->dev);
    if (s->cur == s->old) s->cur = free_buf(s, s->old, s->old);  // keep qold intact
    HIPCHK(hipMemcpyAsync(s->buf[s->cur], q, s->n * sizeof(double), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    s->have_q = true;
    return SC_OK;
}

// ghosts of the current q hold the boundary fill of its interior
extern "C" int sc_get_q(sc_solver *s, double *q)
{
    if (!s || !q) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    DevGuard g(s->dev);
    if (s->cur == s->old) {
        // q and qold share a buffer: fill ghosts of that buffer (harmless for qold)
    }
    HIPCHK(launch_bc(s->p.ndim, s->p.meqn, s->buf[s->cur], s->p.nx[0], s->p.nx[1],
                     s->p.nx[0] + 6, s->p.ndim == 2 ? s->p.nx[1] + 6 : 1, s->p.mthbc, s->st));
    HIPCHK(hipMemcpyAsync(q, s->buf[s->cur], s->n * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return SC_OK;
}

// ----------------------------------------------------------- frame output

extern "C" int sc_output_begin(sc_solver *s)
{
    if (!s) return sc_fail(SC_ERR_ARG, "NULL solver");
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    if (s->out_pending) return sc_fail(SC_ERR_STATE, "the previous frame has not been collected (sc_output_wait)");
    DevGuard g(s->dev);
    if (!s->snap || !s->hsnap || !s->cst || !s->ev_snap || !s->ev_out) {
        output_free(s);
        hipError_t e = hipMalloc(&s->snap, s->n * sizeof(double));
        if (e == hipSuccess) e = hipHostMalloc(&s->hsnap, s->n * sizeof(double), hipHostMallocDefault);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->cst, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_snap, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming);
        if (e != hipSuccess) {
            output_free(s);
            return sc_fail(SC_ERR_HIP, "frame output buffers: %s", hipGetErrorString(e));
        }
    }
    const int LX = s->p.nx[0] + 6, LY = s->p.ndim == 2 ? s->p.nx[1] + 6 : 1;
    double *q = s->buf[s->cur];
    HIPCHK(launch_flush_small(s->p.ndim, s->p.meqn, q, s->p.nx[0], s->p.nx[1], LX, LY, s->st));
    HIPCHK(launch_bc(s->p.ndim, s->p.meqn, q, s->p.nx[0], s->p.nx[1], LX, LY, s->p.mthbc, s->st));
    HIPCHK(hipMemcpyAsync(s->snap, q, s->n * sizeof(double), hipMemcpyDeviceToDevice, s->st));
    HIPCHK(hipEventRecord(s->ev_snap, s->st));
    HIPCHK(hipStreamWaitEvent(s->cst, s->ev_snap, 0));
    HIPCHK(hipMemcpyAsync(s->hsnap, s->snap, s->n * sizeof(double), hipMemcpyDeviceToHost, s->cst));
    HIPCHK(hipEventRecord(s->ev_out, s->cst));
    s->out_pending = true;
    return SC_OK;
}

extern "C" int sc_output_wait(sc_solver *s, double *q)
{
    if (!s || !q) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (!s->out_pending) return sc_fail(SC_ERR_STATE, "no frame in flight (sc_output_begin)");
    DevGuard g(s->dev);
    HIPCHK(hipEventSynchronize(s->ev_out));
    memcpy(q, s->hsnap, s->n * sizeof(double));
    s->out_pending = false;
    return SC_OK;
}

// ---------------------------------------------------------------- restart

namespace {

struct RestartHeader {
    char magic[8];  // "SCRSTRT1"
    int32_t ndim, nx, ny, meqn, mbc, physics;
    double t, dt_next;
    uint64_t n;         // doubles that follow (the sc_get_q layout)
    uint64_t checksum;  // FNV-1a 64 of those bytes
};

uint64_t fnv1a(const void *data, size_t bytes)
{
    const unsigned char *b = (const unsigned char *)data;
    uint64_t h = 1469598103934665603ull;
    for (size_t k = 0; k < bytes; k++) {
        h ^= b[k];
        h *= 1099511628211ull;
    }
    return h;
}

}  // namespace

extern "C" int sc_write_restart(sc_solver *s, const char *path, double t, double dt_next)
{
    if (!s || !path) return sc_fail(SC_ERR_ARG, "NULL argument");
    std::vector<double> q(s->n);
    int rc = sc_get_q(s, q.data());
    if (rc) return rc;
    RestartHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "SCRSTRT1", 8);
    h.ndim = s->p.ndim;
    h.nx = s->p.nx[0];
    h.ny = s->p.ndim == 2 ? s->p.nx[1] : 1;
    h.meqn = s->p.meqn;
    h.mbc = s->p.mbc;
    h.physics = s->p.physics;
    h.t = t;
    h.dt_next = dt_next;
    h.n = s->n;
    h.checksum = fnv1a(q.data(), s->n * sizeof(double));
    std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return sc_fail(SC_ERR_ARG, "cannot open %s for writing", tmp.c_str());
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(q.data(), sizeof(double), s->n, f) == s->n;
    ok = (fclose(f) == 0) && ok;
    // the file appears complete or not at all
    ok = ok && rename(tmp.c_str(), path) == 0;
    if (!ok) {
        remove(tmp.c_str());
        return sc_fail(SC_ERR_ARG, "writing %s failed", path);
    }
    return SC_OK;
}

extern "C" int sc_read_restart(sc_solver *s, const char *path, double *t, double *dt_next)
{
    if (!s || !path || !t || !dt_next) return sc_fail(SC_ERR_ARG, "NULL argument");
    FILE *f = fopen(path, "rb");
    if (!f) return sc_fail(SC_ERR_ARG, "cannot open %s", path);
    RestartHeader h;
    std::vector<double> q(s->n);
    bool ok = fread(&h, sizeof h, 1, f) == 1;
    int ny = s->p.ndim == 2 ? s->p.nx[1] : 1;
    if (ok && memcmp(h.magic, "SCRSTRT1", 8) != 0) {
        fclose(f);
        return sc_fail(SC_ERR_ARG, "%s is not a restart file", path);
    }
    if (ok && (h.ndim != s->p.ndim || h.nx != s->p.nx[0] || h.ny != ny || h.meqn != s->p.meqn ||
               h.mbc != s->p.mbc || h.physics != s->p.physics || h.n != s->n)) {
        fclose(f);
        return sc_fail(SC_ERR_ARG, "%s holds a %dD grid %dx%d, meqn %d, physics %d: not this solver's", path,
                       h.ndim, h.nx, h.ny, h.meqn, h.physics);
    }
    ok = ok && fread(q.data(), sizeof(double), s->n, f) == s->n;
    fclose(f);
    if (!ok) return sc_fail(SC_ERR_ARG, "%s is truncated", path);
    if (fnv1a(q.data(), s->n * sizeof(double)) != h.checksum)
        return sc_fail(SC_ERR_ARG, "%s: checksum mismatch", path);
    int rc = sc_set_q(s, q.data());
    if (rc) return rc;
    *t = h.t;
    *dt_next = h.dt_next;
    return SC_OK;
}

// ------------------------------------------------------------------- aux

extern "C" int sc_set_aux(sc_solver *s, const double *aux)
{
    if (!s || !aux) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (s->p.maux < 1) return sc_fail(SC_ERR_STATE, "maux = 0: the solver has no aux arrays");
    DevGuard g(s->dev);
    if (s->sel.phys == ADV2) {
        // edge velocities aux(:,:,1) (x) and aux(:,:,2) (y), rpn2ad1.f:47
        const int nx = s->p.nx[0], ny = s->p.nx[1];
        const size_t LX = (size_t)nx + 6, LY = (size_t)ny + 6, plane = LX * LY;
        if (!s->aux2 && hipMalloc(&s->aux2, 2 * plane * sizeof(double)) != hipSuccess)
            return sc_fail(SC_ERR_HIP, "aux: %s", hipGetErrorString(hipGetLastError()));
        HIPCHK(hipMemcpyAsync(s->aux2, aux, 2 * plane * sizeof(double), hipMemcpyHostToDevice, s->st));
        auto A = [&](int i, int j, int d) { return aux[(size_t)(i + 2) + LX * (size_t)(j + 2) + (size_t)d * plane]; };
        double gx = 0.0, gy = 0.0;
        for (int i = 1; i <= nx + 1; i++) gx = std::fmax(gx, std::fmax(std::fabs(A(i, 0, 0)), std::fabs(A(i, ny + 1, 0))));
        for (int j = 1; j <= ny + 1; j++) gy = std::fmax(gy, std::fmax(std::fabs(A(0, j, 1)), std::fabs(A(nx + 1, j, 1))));
        s->auxg[0] = gx;
        s->auxg[1] = gy;
        HIPCHK(hipStreamSynchronize(s->st));
        return SC_OK;
    }
    if (s->p.mcapa < 1) return SC_OK;  // no device physics reads aux otherwise
    // the capacity field aux(:,mcapa) / aux(:,:,mcapa), ghost cells included
    size_t LX = (size_t)s->p.nx[0] + 2 * s->p.mbc;
    if (s->p.ndim == 2) LX *= (size_t)s->p.nx[1] + 2 * s->p.mbc;
    if (!s->capa && hipMalloc(&s->capa, LX * sizeof(double)) != hipSuccess)
        return sc_fail(SC_ERR_HIP, "aux: %s", hipGetErrorString(hipGetLastError()));
    HIPCHK(hipMemcpyAsync(s->capa, aux + (size_t)(s->p.mcapa - 1) * LX, LX * sizeof(double), hipMemcpyHostToDevice,
                          s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    return SC_OK;
}

extern "C" int sc_save_qold(sc_solver *s)
{
    if (!s) return sc_fail(SC_ERR_ARG, "NULL solver");
    s->vt_stale = true;
    s->old = s->cur;
    return SC_OK;
}

extern "C" int sc_restore_qold(sc_solver *s)
{
    if (!s) return sc_fail(SC_ERR_ARG, "NULL solver");
    s->cur = s->old;
    return SC_OK;
}

static StageArgs base_args(const sc_solver *s)
{
    StageArgs a;
    memset(&a, 0, sizeof a);
    a.nx = s->p.nx[0];
    a.ny = s->p.ndim == 2 ? s->p.nx[1] : 1;
    a.LX = s->p.nx[0] + 2 * s->p.mbc;
    a.LY = s->p.ndim == 2 ? s->p.nx[1] + 2 * s->p.mbc : 1;
    for (int k = 0; k < 4; k++) a.bc[k] = s->p.mthbc[k];
    a.capa = s->p.mcapa > 0 ? s->capa : nullptr;
    // 2D with a capacity: the constant-tile shortcuts take the CFL speeds of
    // a tile from one cell, which a varying kappa does not allow
    a.full = s->p.ndim == 2 && a.capa ? 1 : 0;
    a.src = s->p.src_term;
    a.aux = s->aux2;
    a.vscale = s->vscale;
    a.xlo = s->p.xlower[0];
    a.dxc = s->p.dx[0];
    a.order = TILE_BANDS;  // time steps set the solver's tile order (enqueue_ts)
    return a;
}

static hipError_t run_stage(sc_solver *s, const StageArgs &a)
{
    if (s->p.ndim == 1) return launch_stage1d(s->sel, 1, a, s->cd, s->cf, s->st);
    if (s->dqx) return launch_sweep2d(s->sel, a, s->dqx, s->cd, s->cf, s->st);
    return launch_stage2d(s->sel, s->tile2d < 0 ? 0 : s->tile2d, a, s->cd, s->cf, s->st);
}

// Autotuner of the 2D tile shape: every candidate computes the same values
// (a tile only changes which workgroup evaluates an interface), so the
// choice is made by timing alone -- two launches of a three-array stage on
// the solver's own q after a warm-up -- and remembered per (physics,
// reconstruction, precision, grid) for the life of the process.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int, int, int>, int> g_tune;

static int tune_tile(sc_solver *s)
{
    if (s->p.ndim != 2 || s->tile2d >= 0 || s->dqx) return SC_OK;
    if (s->p.mcapa > 0) {
        s->tile2d = 0;  // the capacity variant is built for the 16x15 tile
        return SC_OK;
    }
    // one tuning per size class (powers of two per direction), so grids of
    // many sizes (served requests) do not each pay for a tuning
    auto lg = [](int n) { int k = 0; while ((2 << k) <= n) k++; return k; };
    auto key = std::make_tuple(s->sel.phys, s->sel.chr, s->sel.fp32, lg(s->p.nx[0]), lg(s->p.nx[1]));
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            s->tile2d = it->second;
            return SC_OK;
        }
    }
    int best = 0;
    if ((long long)s->p.nx[0] * s->p.nx[1] >= 128 * 128) {
        StageArgs a = base_args(s);
        a.dtdx = 1e-6 / s->p.dx[0];
        a.dtdy = 1e-6 / s->p.dx[1];
        a.qin = s->buf[s->cur];
        a.qold = s->buf[s->cur];
        a.qout = s->buf[free_buf(s, s->cur, s->old)];
        a.op = OP_SSP33_2;
        a.smax = nullptr;
        a.full = 1;  // time the full path, whatever share of the data is constant
        float bestms = 1e30f;
        for (int k = 0; k < NTILE2D; k++) {
            HIPCHK(launch_stage2d(s->sel, k, a, s->cd, s->cf, s->st));
            HIPCHK(hipEventRecord(s->ev0, s->st));
            for (int r = 0; r < 2; r++) HIPCHK(launch_stage2d(s->sel, k, a, s->cd, s->cf, s->st));
            HIPCHK(hipEventRecord(s->ev1, s->st));
            HIPCHK(hipEventSynchronize(s->ev1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
            if (ms < bestms) {
                bestms = ms;
                best = k;
            }
        }
    }
    s->tile2d = best;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best;
    return SC_OK;
}

extern "C" int sc_set_tile(sc_solver *s, int tile)
{
    if (!s) return sc_fail(SC_ERR_ARG, "NULL solver");
    if (tile < -1 || tile >= NTILE2D) return sc_fail(SC_ERR_ARG, "tile must be -1..%d", NTILE2D - 1);
    if (tile > 0 && s->p.ndim == 2 && s->p.mcapa > 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "a 2D grid with a capacity runs on tile 0 (16x15)");
    s->tile2d = tile;
    return SC_OK;
}

extern "C" int sc_get_tile(sc_solver *s, int *tile, int *tx, int *ty)
{
    if (!s || !tile) return sc_fail(SC_ERR_ARG, "NULL argument");
    *tile = s->tile2d;
    int a = 0, b = 0;
    tile2d_shape(s->tile2d < 0 ? 0 : s->tile2d, &a, &b);
    if (tx) *tx = a;
    if (ty) *ty = b;
    return SC_OK;
}

extern "C" int sc_stage_timing(sc_solver *s, int max_launches)
{
    if (!s || max_launches < 0) return sc_fail(SC_ERR_ARG, "bad argument");
    DevGuard g(s->dev);
    s->timing = false;
    s->tev_used = 0;
    if (max_launches == 0) return SC_OK;
    const int need = 2 * max_launches;
    if (need > s->tev_cap) {
        hipEvent_t *t = (hipEvent_t *)realloc(s->tev, need * sizeof(hipEvent_t));
        if (!t) return sc_fail(SC_ERR_NOMEM, "stage timing: %d events", need);
        for (int k = s->tev_cap; k < need; k++) t[k] = nullptr;
        s->tev = t;
        s->tev_cap = need;
    }
    for (int k = 0; k < need; k++)
        if (!s->tev[k]) HIPCHK(hipEventCreate(&s->tev[k]));
    s->timing = true;
    return SC_OK;
}

extern "C" int sc_stage_timing_get(sc_solver *s, double *ms, long long *launches)
{
    if (!s || !ms || !launches) return sc_fail(SC_ERR_ARG, "NULL argument");
    DevGuard g(s->dev);
    double total = 0.0;
    if (s->tev_used > 0) HIPCHK(hipEventSynchronize(s->tev[s->tev_used - 1]));
    for (int k = 0; k + 1 < s->tev_used; k += 2) {
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, s->tev[k], s->tev[k + 1]));
        total += e;
    }
    *ms = total;
    *launches = s->tev_used / 2;
    return SC_OK;
}

// next CFL set, cleared ring-wise: a set that is not cleared before its
// reuse would keep the atomicMax of an earlier step
static int next_slot(sc_solver *s, unsigned long long **slot)
{
    if (s->slot_ctr % NSLOT == 0 && s->slot_ctr > 0)
        HIPCHK(hipMemsetAsync(s->slots, 0, NSLOT * CFL_WORDS * sizeof(unsigned long long), s->st));
    *slot = s->slots + CFL_WORDS * (s->slot_ctr % NSLOT);
    s->slot_ctr++;
    return SC_OK;
}

// The CFL number of the CFL set w: a one-wave kernel writes the set's two
// words to pinned host memory (s->hcfl), unless the step's tile partition
// already did (sent), and the host reads them once the stream has passed it.
static int read_cfl(sc_solver *s, unsigned long long *w, double dt, double *cfl, bool sent = false)
{
    if (!sent) HIPCHK(launch_cfl_to_host(w, s->hcfl, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    double sx = s->hcfl[0], sy = s->hcfl[1];
    if (s->sel.phys == ADV2) {
        // the ghost rows / columns the reference also sweeps (flux2.f90:36,68)
        sx = std::fmax(sx, std::fabs(s->vscale) * s->auxg[0]);
        sy = std::fmax(sy, std::fabs(s->vscale) * s->auxg[1]);
    }
    // cfl = max over interfaces of dtdx*|s| (flux1.f90:126-133), max of the
    // x and y sweeps in 2D (flux2.f90:47,80); dtdx*|s| is monotone in |s|
    double c = (dt / s->p.dx[0]) * sx;
    if (s->p.ndim == 2) c = std::fmax(c, (dt / s->p.dx[1]) * sy);
    *cfl = c;
    return SC_OK;
}

extern "C" int sc_step(sc_solver *s, double dt, double t, double *dq, double *cfl)
{
    (void)t;
    if (!s || !cfl) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (sc_user_bc(&s->p)) return user_bc_refused();
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    if ((s->p.mcapa > 0 && !s->capa) || (s->sel.phys == ADV2 && !s->aux2))
        return sc_fail(SC_ERR_STATE, "no aux uploaded (sc_set_aux)");
    DevGuard g(s->dev);
    StageArgs a = base_args(s);
    double *q = s->buf[s->cur];
    HIPCHK(launch_bc(s->p.ndim, s->p.meqn, q, a.nx, a.ny, a.LX, a.LY, s->p.mthbc, s->st));
    HIPCHK(hipMemsetAsync(s->dq, 0, s->n * sizeof(double), s->st));
    a.qin = q;
    a.qout = s->dq;
    a.op = OP_DQ;
    a.dtdx = dt / s->p.dx[0];
    a.dt = dt;
    a.dtdy = dt / s->p.dx[1];
    int rs = next_slot(s, &a.smax);
    if (rs) return rs;
    HIPCHK(run_stage(s, a));
    int rc = read_cfl(s, a.smax, dt, cfl);
    if (rc) return rc;
    if (dq) {
        HIPCHK(hipMemcpyAsync(dq, s->dq, s->n * sizeof(double), hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
    }
    return SC_OK;
}

// flux1 / flux2 of a whole grid on the caller's q, ghost cells as given
// (the caller's bc has filled them): the x and y interfaces of every cell
// are solved from q with every boundary read from memory (bc type 0), and
// dq(interior) += dt*L(q).  In 2D the reference also sweeps the ghost rows
// j = 0, ny+1 and columns i = 0, nx+1 for the CFL number (flux2.f90:36,68);
// with arbitrary ghost values their speeds are not those of the interior
// rows, so the tile path solves those four slices as 1D lines (the slice
// sweeps already cover them).
extern "C" int sc_flux(sc_solver *s, const double *q, double *dq, double dt, double t, double *cfl)
{
    (void)t;
    if (!s || !q || !dq || !cfl) return sc_fail(SC_ERR_ARG, "NULL argument");
    if ((s->p.mcapa > 0 && !s->capa) || (s->sel.phys == ADV2 && !s->aux2))
        return sc_fail(SC_ERR_STATE, "no aux uploaded (sc_set_aux)");
    if (s->sel.phys == ADV2 && s->p.tperiod != 0.0)
        return sc_fail(SC_ERR_UNSUPPORTED, "2D advection with tperiod set scales its velocities once per step "
                                           "(sc_ts); with sc_flux the caller's b4step scales aux and tperiod is 0");
    if (s->p.ndim == 2 && s->p.mcapa > 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "sc_flux on a 2D grid with a capacity: the ghost rows' CFL would need "
                                           "their capacities; 2D capacity runs take mthbc 1-3");
    DevGuard g(s->dev);
    StageArgs a = base_args(s);
    for (int k = 0; k < 4; k++) a.bc[k] = 0;
    a.src = 0;  // the caller's src has set dq (step1d.f90:26, step2d.f90:31)
    a.vscale = 1.0;
    const int in = free_buf(s, s->cur, s->old);
    HIPCHK(hipMemcpyAsync(s->buf[in], q, s->n * sizeof(double), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(s->dq, 0, s->n * sizeof(double), s->st));
    a.qin = s->buf[in];
    a.qout = s->dq;
    a.op = OP_DQ;
    a.dtdx = dt / s->p.dx[0];
    a.dt = dt;
    a.dtdy = s->p.ndim == 2 ? dt / s->p.dx[1] : 0.0;
    int rs = next_slot(s, &a.smax);
    if (rs) return rs;
    HIPCHK(run_stage(s, a));
    if (s->p.ndim == 2 && !s->dqx && s->sel.phys != ADV2) {
        // the ghost slices: rows 0 and ny+1 (x lines of nx cells), columns 0
        // and nx+1 (y lines of ny cells), each with its mbc ghost cells
        const int nx = s->p.nx[0], ny = s->p.nx[1], M = s->p.meqn;
        const size_t LX = (size_t)nx + 6, LY = (size_t)ny + 6, plane = LX * LY;
        const size_t nsl = 2 * (LX + LY) * M;
        if (!s->gsl && hipMalloc(&s->gsl, 2 * nsl * sizeof(double)) != hipSuccess)
            return sc_fail(SC_ERR_NOMEM, "sc_flux: ghost slices (%zu doubles)", 2 * nsl);
        std::vector<double> h(nsl);
        double *hp = h.data();
        for (int r = 0; r < 2; r++) {
            const size_t j = r == 0 ? 2 : (size_t)ny + 3;  // rows 0, ny+1 in storage
            for (int m = 0; m < M; m++)
                for (size_t i = 0; i < LX; i++) hp[i + m * LX] = q[i + LX * j + m * plane];
            hp += LX * M;
        }
        for (int c = 0; c < 2; c++) {
            const size_t i = c == 0 ? 2 : (size_t)nx + 3;  // columns 0, nx+1
            for (int m = 0; m < M; m++)
                for (size_t j = 0; j < LY; j++) hp[j + m * LY] = q[i + LX * j + m * plane];
            hp += LY * M;
        }
        HIPCHK(hipMemcpyAsync(s->gsl, h.data(), nsl * sizeof(double), hipMemcpyHostToDevice, s->st));
        size_t off = 0;
        for (int k = 0; k < 4; k++) {
            const int ixy = k < 2 ? 1 : 2;
            const int len = ixy == 1 ? nx : ny;
            StageArgs b;
            memset(&b, 0, sizeof b);
            b.qin = s->gsl + off;
            b.qout = s->gsl + nsl + off;
            b.smax = a.smax + (ixy - 1);  // the x or y word of the same CFL set
            b.nx = len;
            b.ny = 1;
            b.LX = len + 6;
            b.LY = 1;
            b.op = OP_DQ;
            b.dtdx = dt / s->p.dx[ixy - 1];
            HIPCHK(launch_stage1d(s->sel, ixy, b, s->cd, s->cf, s->st));
            off += (size_t)(len + 6) * M;
        }
    }
    int rc = read_cfl(s, a.smax, dt, cfl);
    if (rc) return rc;
    std::vector<double> h(s->n);
    HIPCHK(hipMemcpyAsync(h.data(), s->dq, s->n * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    // dq(interior) += dt*L(q); dq's ghost entries stay the caller's
    const size_t LX = (size_t)s->p.nx[0] + 6;
    const size_t LY = s->p.ndim == 2 ? (size_t)s->p.nx[1] + 6 : 1;
    const int ny = s->p.ndim == 2 ? s->p.nx[1] : 1, j0 = s->p.ndim == 2 ? 3 : 0;
    for (int m = 0; m < s->p.meqn; m++)
        for (int j = 0; j < ny; j++) {
            const size_t base = (size_t)m * LX * LY + LX * (size_t)(j + j0) + 3;
            for (int i = 0; i < s->p.nx[0]; i++) dq[base + i] += h[base + i];
        }
    return SC_OK;
}

// The tile flags of stage k of n in a 2D step: the first stage writes the
// flags its second and third stages read for their deep-constant tiles
// (dflag), the last stage those the tile list is built from (tflag).
static void stage_flags(const sc_solver *s, int k, int n, StageArgs &a)
{
    a.tflag = k == 0 ? s->dflag : k == n - 1 ? s->tflag : nullptr;
    a.dflag = k == 1 || k == 2 ? s->dflag : nullptr;
}

// The stage sequence of ts_ssp22/42/33/104 as (op, reads qold, uses register)
struct StagePlan {
    int n;
    int op[10];
};

static StagePlan plan_of(int ti)
{
    StagePlan P;
    switch (ti) {
    case 1: P = {2, {OP_ADD1, OP_SSP22_2}}; break;
    case 2: P = {4, {OP_ADD_THIRD, OP_ADD_THIRD, OP_ADD_THIRD, OP_SSP42_4}}; break;
    case 3: P = {3, {OP_ADD1, OP_SSP33_2, OP_SSP33_3}}; break;
    default:
        P = {10, {OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_SSP104_5, OP_ADD_SIXTH,
                  OP_ADD_SIXTH, OP_ADD_SIXTH, OP_ADD_SIXTH, OP_SSP104_10}};
    }
    return P;
}

// enqueue one time step; returns the CFL word of its last stage
static int enqueue_ts(sc_solver *s, double dt, unsigned long long **slot, bool *sent)
{
    *sent = false;
    StageArgs a = base_args(s);
    a.dtdx = dt / s->p.dx[0];
    a.dt = dt;
    a.dtdy = dt / s->p.dx[1];
    a.qold = s->buf[s->old];
    a.qbin = s->qb;
    a.qbout = s->qb;
    StagePlan P = plan_of(s->p.time_integrator);
    // the tile order of the 2D tile stage (TILE_ADAPTIVE: the tile list of
    // the last step, rebuilt after this step's last stage)
    const bool tiles = s->p.ndim == 2 && !s->dqx;
    int ntx = 0, nty = 0;
    if (tiles) {
        a.order = s->tile_order;
        int tx, ty;
        const int shape = s->tile2d < 0 ? 0 : s->tile2d;
        tile2d_shape(shape, &tx, &ty);
        ntx = (s->p.nx[0] + tx - 1) / tx;
        nty = (s->p.nx[1] + ty - 1) / ty;
        if (a.order == TILE_ADAPTIVE && s->tnf) {
            if (s->tlist_tile != shape) {
                HIPCHK(launch_tile_list_reset(s->tflag, ntx, nty, s->tlist, s->tnf, s->st));
                s->tlist_tile = shape;
            }
            a.tflag = s->tflag;
            a.tlist = s->tlist;
            a.tnf = s->tnf;
        } else {
            a.order = TILE_BANDS;
        }
    }
    int in = s->cur;
    for (int k = 0; k < P.n; k++) {
        int outb = free_buf(s, in, s->old);
        a.qin = s->buf[in];
        a.qout = s->buf[outb];
        a.op = P.op[k];
        a.smax = nullptr;
        if (a.order == TILE_ADAPTIVE) stage_flags(s, k, P.n, a);
        if (k == P.n - 1) {
            int rs = next_slot(s, &a.smax);
            if (rs) return rs;
            *slot = a.smax;
        }
        const bool tm = s->timing && s->tev_used + 2 <= s->tev_cap;
        if (tm) set_launch_events(s->tev[s->tev_used], s->tev[s->tev_used + 1]);
        hipError_t e = run_stage(s, a);
        if (tm) {
            set_launch_events(nullptr, nullptr);
            if (e == hipSuccess) s->tev_used += 2;
        }
        HIPCHK(e);
        in = outb;
    }
    if (a.order == TILE_ADAPTIVE) {
        HIPCHK(launch_tile_partition(s->tflag, ntx, nty, s->tlist, s->tnf, *slot, s->hcfl, s->tcount, s->st));
        *sent = true;
    }
    s->cur = in;
    return SC_OK;
}

// apps/advection/2d/swirl/b4step.f90:26-41 (called by sharpclaw.f90:129 at
// each new step, not on a retake): velocities = aux0 * cos(pi2 (t + dt/2) /
// tperiod), pi2 = 2 pi with pi = 4 atan(1) (swirl/setprob.f)
static void b4step_factor(sc_solver *s, double told, double dt)
{
    if (s->sel.phys != ADV2 || s->p.tperiod == 0.0) return;
    const double pi = 4.0 * std::atan(1.0), pi2 = 2.0 * pi;
    s->vscale = std::cos(pi2 * (told + dt / 2.0) / s->p.tperiod);
}

extern "C" int sc_ts(sc_solver *s, double dt, double told, double *cfl)
{
    if (!s || !cfl) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (sc_user_bc(&s->p)) return user_bc_refused();
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    if ((s->p.mcapa > 0 && !s->capa) || (s->sel.phys == ADV2 && !s->aux2))
        return sc_fail(SC_ERR_STATE, "no aux uploaded (sc_set_aux)");
    DevGuard g(s->dev);
    int rc = tune_tile(s);
    if (rc) return rc;
    if (s->vt_stale) {
        // first attempt of a step (sc_save_qold): the reference's b4step
        b4step_factor(s, told, dt);
        s->vt_stale = false;
    }
    unsigned long long *slot = nullptr;
    bool sent = false;
    rc = enqueue_ts(s, dt, &slot, &sent);
    if (rc) return rc;
    return read_cfl(s, slot, dt, cfl, sent);
}

static int stages_of(int ti)
{
    static const int n[] = {0, 2, 4, 3, 10};
    return n[ti];
}

// The sharpclaw() loop of a 2D grid of at most 16384 tiles kept on the
// device: the stage launches take their arguments from a control block in
// device memory (Res2D), which a one-block control launch after each
// iteration rewrites from the loop's decisions (CFL, dt, accept/reject,
// buffer rotation; the same arithmetic as the host-driven loop) together
// with the adaptive tile list, so the host enqueues a round of iterations
// without waiting for any of them and reads the loop state back once per
// round.  Launches enqueued past the end of the loop return at once.
static int advance2d_resident(sc_solver *s, const sc_run_ctl *c, double tstart, double tend, double *dt_io,
                              sc_run_stats *st)
{
    if (!s->res) {
        if (hipMalloc(&s->res, sizeof(Res2D)) != hipSuccess ||
            hipMalloc(&s->rset, CFL_WORDS * sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc(&s->hres, sizeof(Res2D), hipHostMallocDefault) != hipSuccess)
            return sc_fail(SC_ERR_NOMEM, "device loop: control block allocation failed");
    }
    const int shape = s->tile2d < 0 ? 0 : s->tile2d;
    int tx, ty;
    tile2d_shape(shape, &tx, &ty);
    const int ntx = (s->p.nx[0] + tx - 1) / tx, nty = (s->p.nx[1] + ty - 1) / ty;
    if (s->tlist_tile != shape) {
        HIPCHK(launch_tile_list_reset(s->tflag, ntx, nty, s->tlist, s->tnf, s->st));
        s->tlist_tile = shape;
    }
    Res2D &R = *s->hres;
    memset(&R, 0, sizeof R);
    R.G.nx = s->p.nx[0];
    R.G.ny = s->p.nx[1];
    R.G.LX = R.G.nx + 2 * s->p.mbc;
    R.G.LY = R.G.ny + 2 * s->p.mbc;
    R.G.ntx = ntx;
    R.G.nty = nty;
    R.G.dx = s->p.dx[0];
    R.G.dy = s->p.dx[1];
    R.G.tstart = tstart;
    R.G.tend = tend;
    R.G.ctl = *c;
    int info0 = 0;
    loop_limits(*c, tstart, tend, *dt_io, &R.G.maxn, &info0);
    R.s = initial_state(tstart, *dt_io, s->cur, info0);
    const StagePlan P = plan_of(s->p.time_integrator);
    R.nst = P.n;
    // the stages' fixed arguments (stage2d_res_kernel); buffers, dt and the
    // tile count come from the control block
    StageArgs A[10];
    for (int k = 0; k < P.n; k++) {
        A[k] = base_args(s);
        A[k].order = TILE_ADAPTIVE;
        A[k].ntx = ntx;
        A[k].nty = nty;
        stage_flags(s, k, P.n, A[k]);
        A[k].tlist = s->tlist;
        A[k].tnf = s->tnf;
        A[k].qbin = s->qb;
        A[k].qbout = s->qb;
        A[k].qin = s->buf[0];  // the three buffers, whose roles the control block gives
        A[k].qold = s->buf[1];
        A[k].qout = s->buf[2];
        A[k].op = P.op[k];
        A[k].smax = k == P.n - 1 ? s->rset : nullptr;
    }
    HIPCHK(hipEventRecord(s->ev0, s->st));
    HIPCHK(hipMemcpyAsync(s->res, &R, sizeof R, hipMemcpyHostToDevice, s->st));
    HIPCHK(launch_res2d_ctl(s->res, s->tflag, s->tlist, s->tnf, s->rset, ntx, nty, 0, 1, s->st));  // open the first iteration
    SolveState S = R.s;
    for (;;) {
        // iterations of this round: those left, and no more than reach tend
        // at the current dt (one more for a retake)
        long long k = (long long)R.G.maxn - S.it + (S.go ? 1 : 0);
        if (tend > S.t && S.dt > 0.0) k = std::min(k, (long long)std::ceil((tend - S.t) / S.dt) + 1);
        const int K = (int)std::max(1LL, std::min(64LL, k));
        for (int it = 0; it < K; it++) {
            for (int j = 0; j < P.n; j++) HIPCHK(launch_stage2d_res(s->sel, shape, A[j], s->res, j, s->cd, s->cf, s->st));
            HIPCHK(launch_res2d_ctl(s->res, s->tflag, s->tlist, s->tnf, s->rset, ntx, nty, 1, 1, s->st));
        }
        HIPCHK(hipMemcpyAsync(&R.s, (const char *)s->res + offsetof(Res2D, s), sizeof(SolveState),
                              hipMemcpyDeviceToHost, s->st));
        HIPCHK(hipStreamSynchronize(s->st));
        S = R.s;
        if (S.done) break;
    }
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    s->cur = S.cur;
    s->old = S.old;
    memset(st, 0, sizeof *st);
    st->gpu_ms = ms;
    st->t = S.t;
    st->cfl_maxused = S.cflmax;
    st->cfl_last = S.cfl;
    st->dt_minused = S.dtmin;
    st->dt_maxused = S.dtmax;
    st->dt_last = S.dt;
    st->steps_taken = S.steps;
    st->info = S.info;
    st->rejected = S.rejected;
    st->stage_evals = S.stages;
    *dt_io = S.dt;
    return SC_OK;
}

// The sharpclaw() loop of a 2D grid kept on the device (sc_batch2d.h): the
// CFL, dt update and accept/reject of every step are taken by a control
// kernel, and the host reads the loop state back once per round of K steps
// instead of the CFL once per step.  With use_graph, rounds after the first
// replay one captured hipGraph (measured slower on ROCm 7 for this pattern:
// 67.2 ms against 62.6 ms eager for 200 steps of 1024^2 Euler, so it is off
// by default).  Same arithmetic as the host-driven loop below.
static int advance2d_device(sc_solver *s, const sc_run_ctl *c, double tstart, double tend, double *dt_io,
                            sc_run_stats *st)
{
    if (!s->dr) {
        s->dr = new Batch2DRunner();
        int rc = s->dr->init(s->p, s->st);
        if (!rc) rc = s->dr->reserve(1);
        if (rc) {
            s->dr->destroy();
            delete s->dr;
            s->dr = nullptr;
            return rc;
        }
    }
    Batch2DRunner &R = *s->dr;
    R.set_tile(s->tile2d);
    int tx, ty;
    tile2d_shape(s->tile2d, &tx, &ty);
    Grid2D &G = R.h_grids[0];
    memset(&G, 0, sizeof G);
    G.nx = s->p.nx[0];
    G.ny = s->p.nx[1];
    G.LX = G.nx + 2 * s->p.mbc;
    G.LY = G.ny + 2 * s->p.mbc;
    G.ntx = (G.nx + tx - 1) / tx;
    G.nty = (G.ny + ty - 1) / ty;
    G.dx = s->p.dx[0];
    G.dy = s->p.dx[1];
    G.tstart = tstart;
    G.tend = tend;
    G.ctl = *c;
    int info0 = 0;
    loop_limits(*c, tstart, tend, *dt_io, &G.maxn, &info0);
    R.h_state[0] = initial_state(tstart, *dt_io, s->cur, info0);
    const int K = round_iterations(stages_of(s->p.time_integrator), (long long)G.ntx * G.nty);
    HIPCHK(hipEventRecord(s->ev0, s->st));
    int rc = R.launch_round(1, K, s->mem);
    if (!rc) rc = R.wait();
    while (!rc && !R.h_state[0].done) {
        rc = R.continue_round(1, K, s->mem, s->use_graph);
        if (!rc) rc = R.wait();
    }
    if (rc) return rc;
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    const SolveState &S = R.h_state[0];
    s->cur = S.cur;
    s->old = S.old;
    memset(st, 0, sizeof *st);
    st->gpu_ms = ms;
    st->t = S.t;
    st->cfl_maxused = S.cflmax;
    st->cfl_last = S.cfl;
    st->dt_minused = S.dtmin;
    st->dt_maxused = S.dtmax;
    st->dt_last = S.dt;
    st->steps_taken = S.steps;
    st->info = S.info;
    st->rejected = S.rejected;
    st->stage_evals = S.stages;
    *dt_io = S.dt;
    return SC_OK;
}

// Options the persistent batch kernel runs a 1D grid with: any
// reconstruction, no capacity, aux or source, no tfluct, boundary options 1-3.
static bool loop1d_fits(const sc_solver *s)
{
    const sc_params &p = s->p;
    return p.ndim == 1 && batch_rk(s->sel) >= 0 && p.mcapa == 0 &&
           p.maux == 0 && p.src_term == 0 && p.tfluct_solver == 0 && p.mthbc[0] >= 1 && p.mthbc[0] <= 3 &&
           p.mthbc[1] >= 1 && p.mthbc[1] <= 3;
}

// Workgroup size of the 1D device loop: SC_LOOP_BLOCK, else 128 (1600-cell
// Euler grid, 566 SSP(10,4) steps, a barrier per step: 16.4 ms at 128
// threads in 22 parts, 16.8 ms at 256 in 8, 21.5 ms at 512 in 4; a barrier
// per stage: 24.2 / 24.8 / 28.5 ms; host-driven loop 31.4 ms).
static int loop1d_block(const PhysSel &sel)
{
    if (const char *e = getenv("SC_LOOP_BLOCK")) {
        const int v = atoi(e);
        for (int k = 0; k < NBATCH_BLOCKS; k++)
            if (kBatchBlocks[k] == v && v <= batch_max_block(sel)) return v;
    }
    return 128;
}

// src/sharpclaw.f90:35-235 on the device for a 1D grid: the time loop of
// the persistent batch kernel (sc_batch_kernel.h) with the grid as its one
// solve, split over workgroups (at most MAX_PARTS); no host round trip per
// step.  Without periodic boundaries each workgroup owns a range of cells
// of at most NT - 1 - 6 (stages - 1) (one chunk per stage with the overlap
// its remaining stages read) and the workgroups meet once per step, for the
// CFL; with periodic ones (or SC_LOOP_STAGEWISE=1) they own chunks and meet
// after every stage.  The stage arithmetic, the CFL maximum and the loop's
// decisions are those of the host-driven loop, so the result is bitwise the
// same.
static int advance1d_device(sc_solver *s, const sc_run_ctl *c, double tstart, double tend, double *dt_io,
                            sc_run_stats *st)
{
    sc_solver::Loop1D &L = s->l1;
    if (!L.solve) {
        // one allocation: descriptor, state, sync words, counter, queue
        const size_t bytes = 256 + 256 + 4 * sizeof(unsigned long long) + 64 + MAX_PARTS * sizeof(int) + 256;
        char *m = nullptr;
        HIPCHK(hipMalloc(&m, bytes));
        L.solve = (BatchSolve *)m;
        L.state = (SolveState *)(m + 256);
        L.sync = (unsigned long long *)(m + 512);
        L.next = (unsigned long long *)(m + 512 + 4 * sizeof(unsigned long long));
        L.order = (int *)(m + 512 + 4 * sizeof(unsigned long long) + 64);
        static_assert(sizeof(BatchSolve) <= 256 && sizeof(SolveState) <= 256, "loop descriptor layout");
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, s->dev));
        L.ncu = prop.multiProcessorCount;
        int khz = 0;
        HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->dev));
        L.tick_ms = khz > 0 ? 1.0 / khz : 1e-5;
    }
    const int nt = loop1d_block(s->sel);
    const int mx = s->p.nx[0];
    const int nch = (mx + nt - 2) / (nt - 1);
    const int cap = std::min(MAX_PARTS, L.ncu * batch_occupancy(s->sel, nt));
    static const int kStages[5] = {0, 2, 4, 3, 10};
    const int own = nt - 1 - 6 * (kStages[s->p.time_integrator] - 1);
    const bool whole = s->p.mthbc[0] != 2 && s->p.mthbc[1] != 2 && own >= 16 && !getenv("SC_LOOP_STAGEWISE");
    const int parts = std::max(1, whole ? std::min((mx + own - 1) / own, cap) : std::min(nch, cap));
    const size_t len = (size_t)(mx + 6) * s->p.meqn;
    const size_t pbytes = whole && parts > 1 ? (size_t)parts * 3 * len * sizeof(double) : 0;
    if (pbytes > L.priv_bytes) {
        if (L.priv) HIPCHK(hipFree(L.priv));
        L.priv = nullptr;
        L.priv_bytes = 0;
        HIPCHK(hipMalloc(&L.priv, pbytes));
        L.priv_bytes = pbytes;
        // every value a stage reads is written before it; the clear keeps
        // the rest defined
        HIPCHK(hipMemsetAsync(L.priv, 0, pbytes, s->st));
    }
    int maxn, info0;
    loop_limits(*c, tstart, tend, *dt_io, &maxn, &info0);
    BatchSolve S;
    memset(&S, 0, sizeof S);
    S.off = 0;
    S.mx = mx;
    S.maxn = maxn;
    S.dx = s->p.dx[0];
    S.tstart = tstart;
    S.tend = tend;
    S.ctl = *c;
    SolveState ss = initial_state(tstart, *dt_io, s->cur, info0);
    int ord[MAX_PARTS];
    for (int p = 0; p < parts; p++) ord[p] = queue_entry(0, p, parts);
    BatchArgs A;
    memset(&A, 0, sizeof A);
    A.solves = L.solve;
    A.state = L.state;
    A.order = L.order;
    A.nsolves = parts;
    A.next = L.next;
    A.ends = 1;
    A.base = s->mem;  // buf[0..2] and qb at k * n: the batch layout with len = n
    A.meqn = s->p.meqn;
    A.bc[0] = s->p.mthbc[0];
    A.bc[1] = s->p.mthbc[1];
    A.ti = s->p.time_integrator;
    A.sync = L.sync;
    A.spin_ticks = (long long)(10000.0 / L.tick_ms);
    A.priv = pbytes ? L.priv : nullptr;
    HIPCHK(hipMemcpyAsync(L.solve, &S, sizeof S, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(L.state, &ss, sizeof ss, hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemcpyAsync(L.order, ord, parts * sizeof(int), hipMemcpyHostToDevice, s->st));
    HIPCHK(hipMemsetAsync(L.sync, 0, 4 * sizeof(unsigned long long) + 64, s->st));
    HIPCHK(hipEventRecord(s->ev0, s->st));
    HIPCHK(launch_batch(s->sel, nt, A, s->cd, s->cf, L.ncu, parts, s->st));
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipMemcpyAsync(&ss, L.state, sizeof ss, hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    if (ss.info == INFO_GANG_TIMEOUT)
        return sc_fail(SC_ERR_HIP, "device loop: the grid's %d workgroups did not meet at a barrier", parts);
    s->cur = ss.cur;
    s->old = ss.old;
    memset(st, 0, sizeof *st);
    st->t = ss.t;
    st->cfl_maxused = ss.cflmax;
    st->cfl_last = ss.cfl;
    st->dt_minused = ss.dtmin;
    st->dt_maxused = ss.dtmax;
    st->dt_last = ss.dt;
    st->steps_taken = ss.steps;
    st->info = ss.info;
    st->rejected = ss.rejected;
    st->stage_evals = ss.stages;
    st->gpu_ms = ms;
    *dt_io = ss.dt;
    return SC_OK;
}

// src/sharpclaw.f90:35-235 (qold refreshed at every new step)
extern "C" int sc_advance(sc_solver *s, const sc_run_ctl *c, double tstart, double tend, double *dt_io,
                          sc_run_stats *st)
{
    if (!s || !c || !dt_io || !st) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (sc_user_bc(&s->p)) return user_bc_refused();
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    if ((s->p.mcapa > 0 && !s->capa) || (s->sel.phys == ADV2 && !s->aux2))
        return sc_fail(SC_ERR_STATE, "no aux uploaded (sc_set_aux)");
    DevGuard g(s->dev);
    int trc = tune_tile(s);
    if (trc) return trc;
    if (s->p.ndim == 2) {
        // The device loop saves the per-step CFL readback and the host's
        // turnaround between steps (~13 us at 1024^2); grids above 16384
        // tiles (one control block partitions at most that many) and
        // SC_HOST_LOOP=1 take the host loop, SC_GRAPH=1 the graph-replayed
        // rounds of the batched 2D kernels.
        int tx, ty;
        tile2d_shape(s->tile2d, &tx, &ty);
        long long tiles = (long long)((s->p.nx[0] + tx - 1) / tx) * ((s->p.nx[1] + ty - 1) / ty);
        // 2D advection steps its velocities per step on the host (b4step)
        // (the slice sweeps have no device-loop kernel)
        // (per-launch stage timing brackets launches from the host loop)
        // (a capacity runs on the host loop's stage kernel)
        bool device = s->sel.phys != ADV2 && !s->dqx && !s->timing && s->p.mcapa == 0 &&
                      s->loop_mode != 0;
        if (device && s->use_graph) return advance2d_device(s, c, tstart, tend, dt_io, st);
        if (device && tiles <= 16384 && s->tnf && s->tile_order == TILE_ADAPTIVE)
            return advance2d_resident(s, c, tstart, tend, dt_io, st);
        if (device && s->loop_mode == 1) return advance2d_device(s, c, tstart, tend, dt_io, st);
    } else if (loop1d_fits(s) && !s->timing) {
        // grids of up to MAX_PARTS chunks loop on the device: per step the
        // host-driven loop waits for the CFL (a few microseconds per stage
        // on a grid of a few chunks); larger grids fill the device with one
        // launch per stage instead
        const int nt = loop1d_block(s->sel);
        const int nch = (s->p.nx[0] + nt - 2) / (nt - 1);
        if (s->loop_mode == 1 || (s->loop_mode < 0 && nch <= MAX_PARTS))
            return advance1d_device(s, c, tstart, tend, dt_io, st);
    }
    memset(st, 0, sizeof *st);
    double t = tstart, dt = *dt_io, told = tstart, cflmax = 0.0, cfl = 0.0;
    double dtmin = dt, dtmax = dt;
    int maxn = c->max_steps, info = 0, steps_taken = 0;
    bool retake = false;
    HIPCHK(hipEventRecord(s->ev0, s->st));
    if (c->dt_variable == 0) {
        if (tend < tstart) {
            maxn = 1;
        } else {
            maxn = (int)((tend - tstart + 1e-10) / dt);
            if (std::fabs(maxn * dt - (tend - tstart)) > 1e-5 * (tend - tstart)) {
                info = 2;
                goto done;
            }
        }
    }
    if (c->dt_variable == 1 && c->cfl_desired > c->cfl_max) {
        info = 3;
        goto done;
    }
    for (int it = 1; it <= maxn; it++) {
        if (!retake) {
            told = t;
            if (told + dt > tend && tstart < tend) dt = tend - told;
            b4step_factor(s, told, dt);
            s->old = s->cur;  // qold = q
        }
        retake = false;
        t = told + dt;
        unsigned long long *slot = nullptr;
        bool sent = false;
        int rc = enqueue_ts(s, dt, &slot, &sent);
        if (rc) return rc;
        st->stage_evals += stages_of(s->p.time_integrator);
        rc = read_cfl(s, slot, dt, &cfl, sent);
        if (rc) return rc;
        if (c->dt_variable == 1) {
            if (cfl > 0.0) {
                dt = std::fmin(c->dt_max, dt * c->cfl_desired / cfl);
                dtmin = std::fmin(dt, dtmin);
                dtmax = std::fmax(dt, dtmax);
            } else {
                dt = c->dt_max;
            }
        }
        if (cfl <= c->cfl_max) {
            cflmax = std::fmax(cfl, cflmax);
        } else {
            t = told;
            s->cur = s->old;  // q = qold
            st->rejected++;
            if (c->dt_variable == 1) {
                retake = true;
            } else {
                cflmax = std::fmax(cfl, cflmax);
                break;
            }
        }
        steps_taken++;
        if (t >= tend) break;
    }
    if (c->dt_variable == 1 && t < tend && steps_taken == maxn) info = 11;
    if (c->dt_variable == 0 && cflmax > c->cfl_max) info = 12;
done:
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    st->gpu_ms = ms;
    st->t = t;
    st->cfl_maxused = cflmax;
    st->cfl_last = cfl;
    st->dt_minused = dtmin;
    st->dt_maxused = dtmax;
    st->dt_last = dt;
    st->steps_taken = steps_taken;
    st->info = info;
    *dt_io = dt;
    return SC_OK;
}

extern "C" int sc_sums(sc_solver *s, double *sums)
{
    if (!s || !sums) return sc_fail(SC_ERR_ARG, "NULL argument");
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    DevGuard g(s->dev);
    StageArgs a = base_args(s);
    HIPCHK(launch_sums(s->p.ndim, s->p.meqn, s->buf[s->cur], a.nx, a.ny, a.LX, a.LY, s->partial,
                       s->nblk_sums, s->sums, s->st));
    HIPCHK(hipMemcpyAsync(sums, s->sums, s->p.meqn * sizeof(double), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    double area = s->p.ndim == 2 ? s->p.dx[0] * s->p.dx[1] : s->p.dx[0];
    for (int m = 0; m < s->p.meqn; m++)
        sums[m] = s->p.ndim == 2 ? sums[m] * s->p.dx[0] * s->p.dx[1] : sums[m] * area;
    return SC_OK;
}

extern "C" int sc_time_stage(sc_solver *s, int n, double dt, double *ms_per_stage)
{
    if (!s || !ms_per_stage || n < 1) return sc_fail(SC_ERR_ARG, "bad argument");
    if (!s->have_q) return sc_fail(SC_ERR_STATE, "no q uploaded");
    if ((s->p.mcapa > 0 && !s->capa) || (s->sel.phys == ADV2 && !s->aux2))
        return sc_fail(SC_ERR_STATE, "no aux uploaded (sc_set_aux)");
    DevGuard g(s->dev);
    int trc = tune_tile(s);
    if (trc) return trc;
    StageArgs a = base_args(s);
    a.dtdx = dt / s->p.dx[0];
    a.dt = dt;
    a.dtdy = dt / s->p.dx[1];
    const int q0 = s->cur;
    a.qold = s->buf[q0];
    a.qbin = s->qb;
    a.qbout = s->qb;
    // the time steps' tile order, with the tile list of the last step (not
    // rebuilt here)
    if (s->p.ndim == 2 && !s->dqx && s->tile_order == TILE_ADAPTIVE && s->tnf &&
        s->tlist_tile == (s->tile2d < 0 ? 0 : s->tile2d)) {
        a.order = TILE_ADAPTIVE;
        a.tlist = s->tlist;
        a.tnf = s->tnf;
    }
    StagePlan P = plan_of(s->p.time_integrator);
    // every step restarts from q0, which no stage writes; as in a real time
    // step only the last stage reduces the CFL (into a scratch word)
    auto one_step = [&]() -> hipError_t {
        int in = q0;
        for (int k = 0; k < P.n; k++) {
            int outb = free_buf(s, in, q0);
            a.qin = s->buf[in];
            a.qout = s->buf[outb];
            a.op = P.op[k];
            a.smax = k == P.n - 1 ? s->slots : nullptr;
            hipError_t e = run_stage(s, a);
            if (e != hipSuccess) return e;
            in = outb;
        }
        return hipSuccess;
    };
    HIPCHK(one_step());  // warm-up
    HIPCHK(hipEventRecord(s->ev0, s->st));
    for (int k = 0; k < n; k++) HIPCHK(one_step());
    HIPCHK(hipEventRecord(s->ev1, s->st));
    HIPCHK(hipEventSynchronize(s->ev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
    *ms_per_stage = ms / ((double)n * P.n);
    HIPCHK(hipMemsetAsync(s->slots, 0, NSLOT * CFL_WORDS * sizeof(unsigned long long), s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    s->slot_ctr = 0;
    return SC_OK;
}

extern "C" int sc_copy_bench(int device, size_t nbytes, int reps, double *gbps)
{
    if (!gbps || reps < 1 || nbytes < 16) return sc_fail(SC_ERR_ARG, "bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return sc_fail(SC_ERR_ARG, "device out of range");
    DevGuard g(device);
    size_t n = nbytes / sizeof(double);
    double *a = nullptr, *b = nullptr;
    hipEvent_t e0, e1;
    HIPCHK(hipMalloc(&a, n * sizeof(double)));
    if (hipMalloc(&b, n * sizeof(double)) != hipSuccess) {
        hipFree(a);
        return sc_fail(SC_ERR_NOMEM, "copy bench allocation failed");
    }
    hipError_t e = hipMemset(a, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_copy(b, a, n, 0);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    for (int k = 0; k < reps && e == hipSuccess; k++) e = launch_copy(b, a, n, 0);
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "copy bench: %s", hipGetErrorString(e));
    *gbps = 2.0 * (double)(n * sizeof(double)) * reps / (ms * 1e-3) / 1e9;
    return SC_OK;
}

// ------------------------------------------------------ single routines

namespace {
struct DevBuf {
    double *p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
    hipError_t alloc(size_t n) { return hipMalloc(&p, n * sizeof(double) + 8); }
};
}  // namespace

static int one_call_setup(const sc_params *params)
{
    int rc = sc_check_params(params, false);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (params->device < 0 || params->device >= ndev) return sc_fail(SC_ERR_ARG, "device out of range");
    return SC_OK;
}

#define UP(d, h, n) HIPCHK(hipMemcpy((d).p, (h), (n) * sizeof(double), hipMemcpyHostToDevice))
#define DOWN(h, d, n) HIPCHK(hipMemcpy((h), (d).p, (n) * sizeof(double), hipMemcpyDeviceToHost))

extern "C" int sc_rp(const sc_params *params, int ixy, int mx, const double *ql, const double *qr,
                     double *wave, double *s, double *amdq, double *apdq)
{
    int rc = one_call_setup(params);
    if (rc) return rc;
    if (mx < 1 || (ixy != 1 && ixy != 2) || !ql || !qr || !wave || !s || !amdq || !apdq)
        return sc_fail(SC_ERR_ARG, "bad argument");
    DevGuard g(params->device);
    const int M = params->meqn, W = params->mwaves, L = mx + 2 * params->mbc;
    PhysSel sel{sc_phys_id(params), 0, params->precision == SC_FP32};
    DevBuf dql, dqr, dw, ds, dam, dap;
    HIPCHK(dql.alloc((size_t)L * M)); HIPCHK(dqr.alloc((size_t)L * M));
    HIPCHK(dw.alloc((size_t)L * M * W)); HIPCHK(ds.alloc((size_t)L * W));
    HIPCHK(dam.alloc((size_t)L * M)); HIPCHK(dap.alloc((size_t)L * M));
    UP(dql, ql, (size_t)L * M);
    UP(dqr, qr, (size_t)L * M);
    // entries the reference leaves untouched (i = 1-mbc) keep the caller's values
    UP(dw, wave, (size_t)L * M * W);
    UP(ds, s, (size_t)L * W);
    UP(dam, amdq, (size_t)L * M);
    UP(dap, apdq, (size_t)L * M);
    HIPCHK(launch_rp(sel, ixy, mx, params->mbc, dql.p, dqr.p, dw.p, ds.p, dam.p, dap.p,
                     sc_make_const<double>(params), sc_make_const<float>(params), 0));
    HIPCHK(hipDeviceSynchronize());
    DOWN(wave, dw, (size_t)L * M * W);
    DOWN(s, ds, (size_t)L * W);
    DOWN(amdq, dam, (size_t)L * M);
    DOWN(apdq, dap, (size_t)L * M);
    return SC_OK;
}

extern "C" int sc_evec(const sc_params *params, int ixy, int mx, const double *q, double *evl,
                       double *evr)
{
    int rc = one_call_setup(params);
    if (rc) return rc;
    if (mx < 1 || (ixy != 1 && ixy != 2) || !q || !evl || !evr) return sc_fail(SC_ERR_ARG, "bad argument");
    DevGuard g(params->device);
    const int M = params->meqn, L = mx + 2 * params->mbc;
    PhysSel sel{sc_phys_id(params), 2, params->precision == SC_FP32};
    DevBuf dq, dl, dr;
    HIPCHK(dq.alloc((size_t)L * M)); HIPCHK(dl.alloc((size_t)L * M * M)); HIPCHK(dr.alloc((size_t)L * M * M));
    UP(dq, q, (size_t)L * M);
    UP(dl, evl, (size_t)L * M * M);
    UP(dr, evr, (size_t)L * M * M);
    HIPCHK(launch_evec(sel, ixy, mx, params->mbc, dq.p, dl.p, dr.p, sc_make_const<double>(params),
                       sc_make_const<float>(params), 0));
    HIPCHK(hipDeviceSynchronize());
    DOWN(evl, dl, (size_t)L * M * M);
    DOWN(evr, dr, (size_t)L * M * M);
    return SC_OK;
}

static int weno_common(const sc_params *params, int mx2, const double *q, const double *evl,
                       const double *evr, double *ql, double *qr, int chr)
{
    if (!params) return sc_fail(SC_ERR_ARG, "params is NULL");
    if (params->meqn < 1 || params->meqn > 4) return sc_fail(SC_ERR_UNSUPPORTED, "meqn must be 1..4");
    if (mx2 < 6 || !q || !ql || !qr || (chr == 2 && (!evl || !evr))) return sc_fail(SC_ERR_ARG, "bad argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    DevGuard g(params->device);
    const int M = params->meqn;
    DevBuf dq, dl, dr, dql, dqr;
    HIPCHK(dq.alloc((size_t)mx2 * M)); HIPCHK(dql.alloc((size_t)mx2 * M)); HIPCHK(dqr.alloc((size_t)mx2 * M));
    UP(dq, q, (size_t)mx2 * M);
    UP(dql, ql, (size_t)mx2 * M);
    UP(dqr, qr, (size_t)mx2 * M);
    if (chr == 2) {
        HIPCHK(dl.alloc((size_t)mx2 * M * M)); HIPCHK(dr.alloc((size_t)mx2 * M * M));
        UP(dl, evl, (size_t)mx2 * M * M);
        UP(dr, evr, (size_t)mx2 * M * M);
    }
    HIPCHK(launch_weno(M, chr, params->precision == SC_FP32, mx2, dq.p, dl.p, dr.p, dql.p, dqr.p, 0));
    HIPCHK(hipDeviceSynchronize());
    DOWN(ql, dql, (size_t)mx2 * M);
    DOWN(qr, dqr, (size_t)mx2 * M);
    return SC_OK;
}

extern "C" int sc_weno5(const sc_params *params, int mx2, const double *q, double *ql, double *qr)
{
    return weno_common(params, mx2, q, nullptr, nullptr, ql, qr, 0);
}

extern "C" int sc_weno5_char(const sc_params *params, int mx2, const double *q, const double *evl,
                             const double *evr, double *ql, double *qr)
{
    return weno_common(params, mx2, q, evl, evr, ql, qr, 2);
}

extern "C" int sc_flux1(const sc_params *params, int ixy, int mx, const double *q1d, double *dq1d,
                        double dt, double *cfl)
{
    int rc = one_call_setup(params);
    if (rc) return rc;
    if (mx < 1 || (ixy != 1 && ixy != 2) || (ixy == 2 && params->ndim != 2) || !q1d || !dq1d || !cfl)
        return sc_fail(SC_ERR_ARG, "bad argument");
    if (params->mcapa > 0)
        return sc_fail(SC_ERR_UNSUPPORTED, "sc_flux1 takes no aux array: capacity runs go through sc_set_aux");
    PhysSel sel{sc_phys_id(params), params->char_decomp, params->precision == SC_FP32, params->lim_type};
    if (params->ndim == 2 && sweep2d_path(sel))
        return sc_fail(SC_ERR_UNSUPPORTED, "sc_flux1 on 2D slices reconstructs with weno5 / weno5_char; the "
                                           "other reconstructions run whole 2D stages (sc_step, sc_ts, sc_advance)");
    DevGuard g(params->device);
    const int M = params->meqn, L = mx + 2 * params->mbc;
    DevBuf dq, ddq;
    HIPCHK(dq.alloc((size_t)L * M));
    HIPCHK(ddq.alloc((size_t)L * M));
    unsigned long long *slot = nullptr;
    HIPCHK(hipMalloc(&slot, CFL_WORDS * sizeof(unsigned long long)));
    UP(dq, q1d, (size_t)L * M);
    HIPCHK(hipMemset(ddq.p, 0, (size_t)L * M * sizeof(double)));
    HIPCHK(hipMemset(slot, 0, CFL_WORDS * sizeof(unsigned long long)));
    StageArgs a;
    memset(&a, 0, sizeof a);
    a.qin = dq.p;
    a.qout = ddq.p;
    a.smax = slot;
    a.nx = mx;
    a.ny = 1;
    a.LX = L;
    a.LY = 1;
    a.op = OP_DQ;
    a.dtdx = dt / params->dx[ixy - 1];
    hipError_t e = launch_stage1d(sel, ixy, a, sc_make_const<double>(params), sc_make_const<float>(params), 0);
    std::vector<unsigned long long> w(CFL_WORDS, 0ull);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(w.data(), slot, CFL_WORDS * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    hipFree(slot);
    if (e != hipSuccess) return sc_fail(SC_ERR_HIP, "flux1: %s", hipGetErrorString(e));
    DOWN(dq1d, ddq, (size_t)L * M);
    double sx;
    const unsigned long long wx = cfl_word(w.data(), 0);
    memcpy(&sx, &wx, 8);
    *cfl = a.dtdx * sx;
    return SC_OK;
}
