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
    ++allocs;
    return HYMD_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

double *sync_scratch(size_t n) {
    constexpr int kMaxDevices = 64;
    static thread_local void *buf[kMaxDevices] = {};
    static thread_local size_t cap[kMaxDevices] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("sync_scratch: no current device");
        return nullptr;
    }
    const size_t need = sizeof(double) * (n > 0 ? n : 1);
    if (cap[dev] < need) {
        if (buf[dev]) (void)hipFree(buf[dev]);
        buf[dev] = nullptr;
        cap[dev] = 0;
        hipError_t e = hipMalloc(&buf[dev], need);
        if (e != hipSuccess) {
            buf[dev] = nullptr;
            set_error("hipMalloc(%zu) failed: %s", need, hipGetErrorString(e));
            return nullptr;
        }
        cap[dev] = need;
    }
    return (double *)buf[dev];
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
// Fixed-order reduction of nblocks rows of nv doubles (one workgroup): each
// column is summed by one wavefront, lanes striding over the rows, then a
// shuffle tree, so the order depends on nblocks and nv only.
__device__ void reduce_rows(const double *__restrict__ partials, int nblocks, int nv, double *__restrict__ out) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = wid; j < nv; j += nw) {
        double acc = lane_strided_sum(partials + j, nblocks, nv, lane);
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
        if (lane == 0) out[j] = acc;
    }
}

__global__ void reduce_partials_kernel(const double *__restrict__ partials, int nblocks, int nv,
                                       double *__restrict__ out) {
    reduce_rows(partials, nblocks, nv, out);
}

// workgroup r reduces nblocks[r] rows starting at row row0[r] into out[r * nv ..]
__global__ void reduce_partials_batched_kernel(const double *__restrict__ partials, const int *__restrict__ nblocks,
                                               const int64_t *__restrict__ row0, int nv, double *__restrict__ out) {
    const int r = blockIdx.x;
    reduce_rows(partials + row0[r] * nv, nblocks[r], nv, out + (int64_t)r * nv);
}

__global__ void zero_words_kernel(uint32_t *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

int zero_words(void *ptr, int64_t nwords, hipStream_t s) {
    if (nwords > 0) zero_words_kernel<<<grid_for(nwords), kBlock, 0, s>>>((uint32_t *)ptr, nwords);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int reduce_partials(const double *partials, int nblocks, int nv, double *out, hipStream_t s) {
    reduce_partials_kernel<<<1, kBlock, 0, s>>>(partials, nblocks, nv, out);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int reduce_partials_batched(const double *partials, const int *nblocks, const int64_t *row0, int nv, double *out,
                            int nbatch, hipStream_t s) {
    if (nbatch > 0) reduce_partials_batched_kernel<<<nbatch, kBlock, 0, s>>>(partials, nblocks, row0, nv, out);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

}  // namespace hymd

extern "C" const char *hymd_last_error(void) { return hymd::g_err; }
extern "C" int hymd_abi_version(void) { return 1; }
