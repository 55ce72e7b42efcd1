// Error state, scratch buffers and the deterministic final reduction.
#include <cstdarg>
#include <cstring>

#include "common.h"

namespace hymd {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int DevBuf::reserve(size_t need) {
    if (need <= bytes) return HYMD_OK;
    release();
    if (need == 0) return HYMD_OK;
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(%zu) failed: %s", need, hipGetErrorString(e));
        return HYMD_ERR_NOMEM;
    }
    bytes = need;
    return HYMD_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int HostBuf::reserve(size_t need) {
    if (need <= bytes) return HYMD_OK;
    release();
    hipError_t e = hipHostMalloc(&ptr, need, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipHostMalloc(%zu) failed: %s", need, hipGetErrorString(e));
        return HYMD_ERR_NOMEM;
    }
    bytes = need;
    return HYMD_OK;
}

void HostBuf::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// One block sums `nblocks` rows of `nv` doubles: thread j accumulates column
// entries with stride, then a fixed tree combines the threads.
__global__ void reduce_partials_kernel(const double *__restrict__ partials, int nblocks, int nv,
                                       double *__restrict__ out) {
    __shared__ double red[kBlock];
    for (int j = 0; j < nv; ++j) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[(int64_t)b * nv + j];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[j] = red[0];
        __syncthreads();
    }
}

int reduce_partials(const double *partials, int nblocks, int nv, double *out, hipStream_t s) {
    reduce_partials_kernel<<<1, kBlock, 0, s>>>(partials, nblocks, nv, out);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

}  // namespace hymd

extern "C" const char *hymd_last_error(void) { return hymd::g_err; }
extern "C" int hymd_abi_version(void) { return 1; }
