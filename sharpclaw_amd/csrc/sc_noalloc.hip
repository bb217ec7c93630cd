// sc_step_noalloc: the step() of sc_step on arrays the caller already holds
// on the device, for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The only memory beyond the caller's q, dq and CFL double is one CFL set
// (the x and y max|s| words with their CFL_SUB copies, 8 KiB) in the static
// device array `scratch` below.  Each device holds its own copy of it until
// the process ends; its address on a device is taken once, with that device
// current, by hipGetSymbolAddress, and only that address is passed to
// kernels.  Every call clears the set with a fill kernel before the stage
// writes it, so a call never sees an earlier call's maxima; calls that
// overlap on one device (several host threads, or streams that run
// concurrently) share the set and race.  The CFL number is written to a
// device double by a one-thread kernel, so the whole call is stream-ordered
// work with no host synchronisation and no memset node.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "sc_internal.h"
#include "sc_stage.h"

using namespace sc;

// one CFL set per device; host code uses only the address that
// hipGetSymbolAddress returns for it (scratch_of)
static __device__ __align__(256) unsigned long long scratch[CFL_WORDS];

namespace {

constexpr int kMaxDevices = 64;

std::mutex g_mu;
unsigned long long *g_scratch[kMaxDevices];  // address of `scratch` per device

__global__ void clear_cfl_set(unsigned long long *set)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < CFL_WORDS) set[k] = 0ull;
}

__global__ void zero_fill(double *p, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

// cfl = max over interfaces of dtdx*|s| (flux1.f90:126-133), the max of the
// x and y sweeps in 2D (flux2.f90:47,80)
__global__ void cfl_from_set(const unsigned long long *set, double dtdx, double dtdy, int ndim, double *cfl)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double sx = __longlong_as_double((long long)cfl_word(set, 0));
    double c = dtdx * sx;
    if (ndim == 2) c = fmax(c, dtdy * __longlong_as_double((long long)cfl_word(set, 1)));
    *cfl = c;
}

struct DeviceScope {
    int prev = -1;
    hipError_t err;
    explicit DeviceScope(int d)
    {
        (void)hipGetDevice(&prev);
        err = hipSetDevice(d);
    }
    ~DeviceScope()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

// the device address of `scratch` on the current device `dev`
int scratch_of(int dev, unsigned long long **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return sc_fail(SC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch[dev] = (unsigned long long *)p;
    }
    *out = g_scratch[dev];
    return SC_OK;
}

}  // namespace

#define NA_CHK(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return sc_fail(SC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

extern "C" int sc_step_noalloc(const sc_params *params, double *q, double *dq, double dt, double t,
                               double *cfl, void *stream)
{
    (void)t;
    if (!params || !q || !dq || !cfl) return sc_fail(SC_ERR_ARG, "NULL argument");
    int rc = sc_check_params(params, true);
    if (rc) return rc;
    if (sc_user_bc(params))
        return sc_fail(SC_ERR_UNSUPPORTED, "mthbc 0 (user boundary conditions): step with sc_flux");
    const int phys = sc_phys_id(params);
    PhysSel sel;
    sel.phys = phys;
    sel.chr = params->char_decomp;
    sel.lim = params->lim_type;
    sel.fp32 = params->precision == SC_FP32;
    if (params->ndim == 2 && sweep2d_path(sel))
        return sc_fail(SC_ERR_UNSUPPORTED,
                       "sc_step_noalloc: lim_type %d / char_decomp %d on a 2D grid runs as slice sweeps, "
                       "which need a grid-sized buffer (use sc_step)", params->lim_type, params->char_decomp);
    if (params->mcapa > 0 || phys == ADV2)
        return sc_fail(SC_ERR_UNSUPPORTED, "sc_step_noalloc: options that read aux (capacity, 2D advection) "
                                           "need sc_set_aux (use sc_step)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return sc_fail(SC_ERR_HIP, "no HIP device available");
    if (params->device < 0 || params->device >= ndev || params->device >= kMaxDevices)
        return sc_fail(SC_ERR_ARG, "device %d out of range (%d devices)", params->device, ndev);
    DeviceScope g(params->device);
    if (g.err != hipSuccess)
        return sc_fail(SC_ERR_HIP, "hipSetDevice(%d): %s", params->device, hipGetErrorString(g.err));
    unsigned long long *set = nullptr;
    rc = scratch_of(params->device, &set);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;

    const int nx = params->nx[0], ny = params->ndim == 2 ? params->nx[1] : 1;
    const int LX = nx + 2 * params->mbc, LY = params->ndim == 2 ? ny + 2 * params->mbc : 1;
    const size_t n = (size_t)LX * LY * params->meqn;
    const double dx = params->dx[0], dy = params->ndim == 2 ? params->dx[1] : 1.0;

    StageArgs a;
    std::memset(&a, 0, sizeof a);
    a.nx = nx;
    a.ny = ny;
    a.LX = LX;
    a.LY = LY;
    for (int k = 0; k < 4; k++) a.bc[k] = params->mthbc[k];
    a.src = params->src_term;
    a.vscale = 1.0;
    a.xlo = params->xlower[0];
    a.dxc = dx;
    a.qin = q;
    a.qout = dq;
    a.op = OP_DQ;
    a.dtdx = dt / dx;
    a.dtdy = dt / dy;
    a.dt = dt;
    a.smax = set;
    const PhysConst<double> cd = sc_make_const<double>(params);
    const PhysConst<float> cf = sc_make_const<float>(params);

    NA_CHK(launch_bc(params->ndim, params->meqn, q, nx, ny, LX, LY, params->mthbc, st));
    hipLaunchKernelGGL(clear_cfl_set, dim3((CFL_WORDS + 255) / 256), dim3(256), 0, st, set);
    NA_CHK(hipGetLastError());
    const size_t nblk = std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(zero_fill, dim3((unsigned)nblk), dim3(256), 0, st, dq, n);
    NA_CHK(hipGetLastError());
    if (params->ndim == 1)
        NA_CHK(launch_stage1d(sel, 1, a, cd, cf, st));
    else
        NA_CHK(launch_stage2d(sel, 0, a, cd, cf, st));
    hipLaunchKernelGGL(cfl_from_set, dim3(1), dim3(64), 0, st, (const unsigned long long *)set, a.dtdx, a.dtdy,
                       params->ndim, cfl);
    NA_CHK(hipGetLastError());
    return SC_OK;
}
