// Exclusive prefix sum of int32 counts on the device: tile-local scans and
// tile totals (kScanTile items per workgroup), then the tile offsets added
// (up to kScanFusedMaxTiles tiles each workgroup sums the totals before its
// tile itself; beyond, a one-workgroup scan of the totals comes first).  out has n + 1 entries (out[n] =
// total).  Optionally copies the result into `cursor` and zeroes the input.
#include "common.h"

namespace hymd {

constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

int64_t scan_tiles(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

__device__ int block_exclusive_scan(int v, int *sh, int &total) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < (int)blockDim.x; o <<= 1) {
        int add = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    total = sh[blockDim.x - 1];
    const int incl = sh[threadIdx.x];
    __syncthreads();
    return incl - v;
}

__global__ void scan_tiles_kernel(int *__restrict__ in, int *__restrict__ out, int64_t n,
                                  int *__restrict__ tile_sums, int zero_input) {
    __shared__ int sh[kBlock];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems], s = 0;
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < n) ? in[base + j] : 0;
        s += v[j];
    }
    if (zero_input)
        for (int j = 0; j < kScanItems; ++j)
            if (base + j < n) in[base + j] = 0;
    int total;
    int run = block_exclusive_scan(s, sh, total);
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < n) out[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ void scan_sums_kernel(int *__restrict__ sums, int ntiles, int *__restrict__ grand_total) {
    __shared__ int sh[kBlock];
    const int per = (ntiles + blockDim.x - 1) / blockDim.x;
    const int b = threadIdx.x * per;
    int s = 0;
    for (int j = 0; j < per; ++j)
        if (b + j < ntiles) s += sums[b + j];
    int total;
    int run = block_exclusive_scan(s, sh, total);
    for (int j = 0; j < per; ++j)
        if (b + j < ntiles) {
            const int v = sums[b + j];
            sums[b + j] = run;
            run += v;
        }
    if (threadIdx.x == 0) *grand_total = total;
}

__global__ void scan_add_kernel(int *__restrict__ out, int64_t n, const int *__restrict__ sums,
                                const int *__restrict__ grand_total, int *__restrict__ cursor) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int v = out[i] + sums[i / kScanTile];
        out[i] = v;
        if (cursor) cursor[i] = v;
    }
    if (i == n) out[n] = *grand_total;
}

// Phases two and three in one pass for short scans: a workgroup of kBlock
// outputs lies in one tile t (kScanTile is a multiple of kBlock) and sums the
// tile totals before t itself, in a fixed order (lanes over tiles, then a
// shuffle tree); the workgroup holding index n sums all of them for out[n].
constexpr int kScanFusedMaxTiles = 1024;

__device__ __forceinline__ int block_sum(int v, int *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    return t;
}

__global__ void __launch_bounds__(kBlock) scan_add_fused_kernel(int *__restrict__ out, int64_t n,
                                                                 const int *__restrict__ tile_sums, int ntiles,
                                                                 int *__restrict__ cursor) {
    __shared__ int red[kBlock / 64];
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x;
    const int t = (int)(i0 / kScanTile);
    const bool has_end = i0 <= n && n < i0 + blockDim.x;
    int pre = 0, all = 0;
    for (int j = threadIdx.x; j < ntiles; j += blockDim.x) {
        const int v = tile_sums[j];
        if (j < t) pre += v;
        all += v;
    }
    pre = block_sum(pre, red);
    all = has_end ? block_sum(all, red) : 0;
    const int64_t i = i0 + threadIdx.x;
    if (i < n) {
        const int v = out[i] + pre;
        out[i] = v;
        if (cursor) cursor[i] = v;
    }
    if (i == n) out[n] = all;
}

// tiles: scratch of scan_tiles(n) + 1 ints
int exclusive_scan(int *in, int *out, int64_t n, int *tiles, int *cursor, bool zero_input, hipStream_t s) {
    const int64_t nt = scan_tiles(n);
    HYMD_REQUIRE(nt <= (int64_t)kBlock * 1024, "scan too long (%lld items)", (long long)n);
    if (n > 0) scan_tiles_kernel<<<(unsigned)nt, kBlock, 0, s>>>(in, out, n, tiles, zero_input ? 1 : 0);
    if (nt <= kScanFusedMaxTiles) {
        scan_add_fused_kernel<<<grid_for(n + 1), kBlock, 0, s>>>(out, n, tiles, (int)nt, cursor);
    } else {
        scan_sums_kernel<<<1, kBlock, 0, s>>>(tiles, (int)nt, tiles + nt);
        scan_add_kernel<<<grid_for(n + 1), kBlock, 0, s>>>(out, n, tiles, tiles + nt, cursor);
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

}  // namespace hymd
