// Negative controls of the CPU emulation: each case runs a small kernel with
// one deliberate fault and must be reported (the sanitizer or the
// emulation's deadlock check aborts the process).  They show that the
// checks the scenarios rely on are live.
//
// Usage: emu_selftest heap|lds|wave|ok
#include <cstdio>
#include <cstring>

#include "hip/hip_runtime.h"

// writes one element past the end of its n-element array
__global__ void heap_overflow(double *p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i <= n) p[i] = 1.0;
}

// indexes a 64-element workgroup array with a 128-thread workgroup
__global__ void lds_overflow(double *out)
{
    __shared__ double a[64];
    a[threadIdx.x] = (double)threadIdx.x;
    __syncthreads();
    out[threadIdx.x] = a[threadIdx.x / 2];
}

// a shuffle, reached by half of a wave, that reads the other half
__global__ void divergent_shuffle(double *out)
{
    double v = (double)threadIdx.x;
    if (threadIdx.x < 32) v = __shfl_xor(v, 32, 64);
    out[threadIdx.x] = v;
}

// well-formed: a workgroup reduction with shuffles and a barrier
__global__ void reduce(const double *in, double *out)
{
    __shared__ double part[4];
    double v = in[blockIdx.x * 256 + threadIdx.x];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

int main(int argc, char **argv)
{
    const char *c = argc > 1 ? argv[1] : "ok";
    double *d = nullptr, *o = nullptr;
    if (!std::strcmp(c, "heap")) {
        (void)hipMalloc(&d, 100 * sizeof(double));
        hipLaunchKernelGGL(heap_overflow, dim3(1), dim3(128), 0, nullptr, d, 100);
    } else if (!std::strcmp(c, "lds")) {
        (void)hipMalloc(&o, 128 * sizeof(double));
        hipLaunchKernelGGL(lds_overflow, dim3(1), dim3(128), 0, nullptr, o);
    } else if (!std::strcmp(c, "wave")) {
        (void)hipMalloc(&o, 64 * sizeof(double));
        hipLaunchKernelGGL(divergent_shuffle, dim3(1), dim3(64), 0, nullptr, o);
    } else {
        const int nb = 40;
        (void)hipMalloc(&d, nb * 256 * sizeof(double));
        (void)hipMalloc(&o, nb * sizeof(double));
        double h[nb * 256];
        for (int k = 0; k < nb * 256; k++) h[k] = (double)(k % 1000);
        (void)hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
        hipLaunchKernelGGL(reduce, dim3(nb), dim3(256), 0, nullptr, d, o);
        double r[nb];
        (void)hipMemcpy(r, o, sizeof(r), hipMemcpyDeviceToHost);
        for (int b = 0; b < nb; b++) {
            double s = 0.0;
            for (int k = 0; k < 256; k++) s += h[b * 256 + k];
            if (r[b] != s) {
                std::printf("reduce: block %d gave %g, expected %g\n", b, r[b], s);
                return 1;
            }
        }
    }
    (void)hipFree(d);
    (void)hipFree(o);
    std::printf("selftest %s completed\n", c);
    return 0;
}
