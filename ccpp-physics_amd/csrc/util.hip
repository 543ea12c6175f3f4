// Small fill / normalisation kernels.  Buffers are cleared with kernels (not
// hipMemsetAsync) so the run entries stay safe to record into a HIP graph.
#include "kernels.h"

namespace ccpprad {

__global__ void fill_f64_kernel(double *p, size_t n, double v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_i32_kernel(int *p, size_t n, int v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u8_kernel(unsigned char *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

static unsigned nblocks(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b ? b : 1);
}

void launch_fill_f64(double *p, size_t n, double v, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(fill_f64_kernel, dim3(nblocks(n)), dim3(256), 0, st, p, n, v);
}
void launch_fill_i32(int *p, size_t n, int v, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(fill_i32_kernel, dim3(nblocks(n)), dim3(256), 0, st, p, n, v);
}
void launch_fill_u8(unsigned char *p, size_t n, hipStream_t st)
{
    if (n) hipLaunchKernelGGL(fill_u8_kernel, dim3(nblocks(n)), dim3(256), 0, st, p, n);
}

// toa_src * solcon / sum(toa_src) (rrtmgp_sw_main.F90:355-357): the same for
// every column because the solar source does not vary by column.  FP64 sum.
__global__ void toa_norm_kernel(const double *solar, int ngpt, double solcon, double *o64, float *o32)
{
    __shared__ double s;
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int g = 0; g < ngpt; ++g) t += solar[g];
        s = t;
    }
    __syncthreads();
    for (int g = threadIdx.x; g < ngpt; g += blockDim.x) {
        const double v = solar[g] * solcon / s;
        o64[g] = v;
        o32[g] = (float)v;
    }
}

void launch_toa_norm(const double *solar, int ngpt, double solcon, double *o64, float *o32, hipStream_t st)
{
    hipLaunchKernelGGL(toa_norm_kernel, dim3(1), dim3(128), 0, st, solar, ngpt, solcon, o64, o32);
}

}  // namespace ccpprad
