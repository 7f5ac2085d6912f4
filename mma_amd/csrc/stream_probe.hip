// Bandwidth probes for the benchmark, not part of the product libraries: a streaming copy, a
// read-only pass and a write-only pass over device memory in 16-byte lanes, grid-stride, four
// accesses in flight per lane.  Built into mma_amd/lib/libmmaprobe.so, which only bench.py
// loads: the rates they reach are the practical ceilings the propagation kernel's traffic is
// compared with.
#include <hip/hip_runtime.h>

__global__ void probe_copy_kernel(const double2 *__restrict__ src, double2 *__restrict__ dst, long long n)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

// Reads every element and keeps a per-lane sum; the sum is stored only if it is NaN (never, for
// finite data), so the loads stay while almost nothing is written.
__global__ void probe_read_kernel(const double2 *__restrict__ src, long long n, double *sink)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        acc += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (d.x + d.y);
    }
    for (; i < n; i += stride) acc += src[i].x + src[i].y;
    if (acc != acc) sink[0] = acc;
}

__global__ void probe_write_kernel(double2 *__restrict__ dst, long long n, double v)
{
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const double2 w = make_double2(v, v);
    for (; i + 3 * stride < n; i += 4 * stride) {
        dst[i] = w;
        dst[i + stride] = w;
        dst[i + 2 * stride] = w;
        dst[i + 3 * stride] = w;
    }
    for (; i < n; i += stride) dst[i] = w;
}

extern "C" {

// kind 0: copy src -> dst; 1: read src (dst receives nothing unless src holds a NaN: then one
// double at dst); 2: write dst.  n_bytes a positive multiple of 16, 1..64 workgroups of 256
// lanes per compute unit.  Queues the pass on `stream` and returns a hipError_t code.
int probe_stream(int kind, const double *src, double *dst, long long n_bytes, int groups_per_cu, void *stream)
{
    if (n_bytes < 16 || n_bytes % 16 != 0 || groups_per_cu < 1 || groups_per_cu > 64 || kind < 0 || kind > 2 ||
        (kind != 2 && !src) || !dst)
        return (int)hipErrorInvalidValue;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return (int)e;
    const dim3 grid(cus * groups_per_cu), block(256);
    const long long n = n_bytes / 16;
    hipStream_t s = (hipStream_t)stream;
    if (kind == 0)
        hipLaunchKernelGGL(probe_copy_kernel, grid, block, 0, s, (const double2 *)src, (double2 *)dst, n);
    else if (kind == 1)
        hipLaunchKernelGGL(probe_read_kernel, grid, block, 0, s, (const double2 *)src, n, dst);
    else
        hipLaunchKernelGGL(probe_write_kernel, grid, block, 0, s, (double2 *)dst, n, 1.0);
    return (int)hipGetLastError();
}

} // extern "C"
