// vc_calib.hip -- a plain streaming copy of a known size, used only to
// calibrate rocprofv3's FETCH_SIZE / WRITE_SIZE counters against known HBM
// traffic before they are applied to the solver kernel.  Not part of the
// solver's C ABI.
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ void calib_copy_kernel(const int4 *__restrict__ src, int4 *__restrict__ dst, size_t n16)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n16; i += stride) dst[i] = src[i];
}

// Copies `bytes` (rounded down to 16) between two fresh device buffers
// `reps` times; returns 0 on success.
extern "C" int vcprobe_copy(size_t bytes, int reps)
{
    size_t n16 = bytes / 16;
    int4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n16 * 16) != hipSuccess || hipMalloc(&b, n16 * 16) != hipSuccess) return 1;
    hipMemset(a, 1, n16 * 16);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(calib_copy_kernel, dim3(4096), dim3(256), 0, 0, a, b, n16);
    int rc = hipDeviceSynchronize() == hipSuccess ? 0 : 2;
    hipFree(a);
    hipFree(b);
    return rc;
}
