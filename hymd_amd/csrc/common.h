// Shared helpers of the HIP hot path (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "hymd_amd.h"

namespace hymd {

void set_error(const char *fmt, ...);

#define HYMD_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            hymd::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return HYMD_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

#define HYMD_REQUIRE(cond, ...)          \
    do {                                 \
        if (!(cond)) {                   \
            hymd::set_error(__VA_ARGS__); \
            return HYMD_ERR_INVALID;     \
        }                                \
    } while (0)

#define HYMD_TRY(expr)           \
    do {                         \
        int s_ = (expr);         \
        if (s_ != HYMD_OK) return s_; \
    } while (0)

constexpr int kBlock = 256;

inline int grid_for(int64_t n, int block = kBlock, int cap = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ---------------------------------------------------------------------------
// Cloud-in-cell window (pmesh CIC, node i at x = i*h).  Evaluated in float64
// with contraction disabled so that the CPU fixed-point replica computes the
// same bits: g = x*inv_h; i = floor(g); f = g - i; w0 = 1-f; w1 = f.
// ---------------------------------------------------------------------------
struct CicAxis {
    int i0, i1;
    double w0, w1;
};

__device__ __forceinline__ CicAxis cic_axis(double x, double inv_h, int M) {
#pragma clang fp contract(off)
    CicAxis c;
    double g = x * inv_h;
    double fl = floor(g);
    double f = g - fl;
    long long ii = (long long)fl;
    ii %= M;
    if (ii < 0) ii += M;
    c.i0 = (int)ii;
    c.i1 = (c.i0 + 1 == M) ? 0 : c.i0 + 1;
    c.w0 = 1.0 - f;
    c.w1 = f;
    return c;
}

// mesh geometry passed by value to kernels
struct Mesh {
    int mx, my, mz, zh;     // real edges, zh = mz/2+1
    int64_t g, gc;          // real cells, complex cells
    double inv_h[3];        // M_d / L_d
    double box[3];
};

// complex helpers (float2 = re, im)
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by +i (s>0) or -i (s<0)
__device__ __forceinline__ float2 cmul_i(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// Deterministic block reduction of doubles (fixed order: a shuffle tree
// inside each 64-wide wavefront, then the wavefront sums in order).  All
// kBlock threads of the block must call it.
template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out_block) {
    constexpr int kWaves = kBlock / 64;
    __shared__ double red[NV][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j = 0; j < NV; ++j) {
        double s = v[j];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) {
            double t = 0.0;
            for (int w = 0; w < kWaves; ++w) t += red[j][w];
            out_block[j] = t;
        }
}

// Reduce `nblocks` partial rows of NV doubles (row-major) in fixed order into
// out[NV] with a single block; launch with kBlock threads.
int reduce_partials(const double *partials, int nblocks, int nv, double *out, hipStream_t s);

// Exclusive scan of n int32 counts into out[0..n] (out[n] = total); tiles is
// scratch of scan_tiles(n) + 1 ints; cursor (optional) receives a copy of
// out[0..n-1]; zero_input clears the counts after reading them.
int64_t scan_tiles(int64_t n);
int exclusive_scan(int *in, int *out, int64_t n, int *tiles, int *cursor, bool zero_input, hipStream_t s);

// Scratch growing buffer owned by the library (per handle or per thread).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    void release();
    ~DevBuf() { release(); }
};

// Term arrays held by a persistent topology (device pointers).
struct TopoTerms {
    int64_t n, m2, m3;
    const int32_t *a2, *b2, *a3, *b3, *c3;
    const double *r0, *k2, *t0, *k3;
};
TopoTerms topology_terms(const hymd_topology *t);

// compute_field_force at the particle set sorted by the last update_field on
// this handle (same positions and types)
int field_readout_last(hymd_pm *pm, int n_types, const float *force_mesh, float *force);

// Pinned host staging for small results.
struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    void release();
    ~HostBuf() { release(); }
};

}  // namespace hymd
