// Device pieces shared by the single-mesh field path and the batched
// replica path, so that both compute identical bits.
#pragma once
#include <algorithm>

#include "common.h"

namespace hymd {

// Scatter of particles into key-sorted order: keys[i] = (key, rank of i
// among the particles of that key), start = exclusive scan of the key
// counts; perm[pos] = i, xs/ts/ws the sorted positions, types (0 when
// types == NULL) and weights (if ws != NULL).
// Particle i (key (row, rank), key.x < 0: left out) goes to slot start[row] +
// rank: perm[slot] = i, rec[slot] = (x, y, z, type bits; type 0 without
// types), ws[slot] = w[i] when ws is given.
int launch_key_scatter(const float *x, const int32_t *types, const float *w, int64_t n, const int2 *keys,
                       const int *start, int *perm, float4 *rec, float *ws, hipStream_t s);

// Block-local counting sort for the field readout: thread k holds item k (valid
// for a prefix of the block's threads) with bucket bk in [0, nbk), nbk <=
// kSortBuckets + 1; returns the block-local index of the item thread k takes
// in bucket order (-1 for the invalid threads).  The order within a bucket
// is that of the LDS atomics.  Every thread of the block calls it (it holds
// barriers); cnt has kSortBuckets + 1 ints and order blockDim.x ints of LDS.
constexpr int kSortBuckets = 256;

__device__ __forceinline__ int block_bucket_order(bool valid, int bk, int nbk, int *cnt, int *order) {
    __syncthreads();  // the previous call's reads of cnt[] and order[] are done
    for (int k = threadIdx.x; k < nbk; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const int rank = valid ? atomicAdd(&cnt[bk], 1) : 0;
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the counts, one wavefront
        const int lane = threadIdx.x;
        constexpr int kPer = (kSortBuckets + 1 + 63) / 64;
        int v[kPer], sum = 0;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = lane * kPer + u;
            v[u] = k < nbk ? cnt[k] : 0;
            sum += v[u];
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int u = 0; u < kPer; ++u) {
            const int k = lane * kPer + u;
            if (k < nbk) cnt[k] = run;
            run += v[u];
        }
    }
    __syncthreads();
    if (valid) order[cnt[bk] + rank] = threadIdx.x;
    __syncthreads();
    return valid ? order[threadIdx.x] : -1;
}

// z line (32 floats, one 128-byte cache line, >> zshift) of a position
__device__ __forceinline__ int z_line(float z, double inv_h, int mz, int zshift) {
    return min(max((int)(z * (float)inv_h), 0), mz - 1) >> (5 + zshift);
}

struct PaintArgs {
    int n_types;
    double scale;  // 2^HYMD_FIXED_SHIFT
    double mass[HYMD_MAX_TYPES];
};

// Fixed-point CIC deposits of one particle into one mesh row: the corners
// (a, b, 0) and (a, b, 1) land at z = iz and iz + 1 of row (ix + a, iy + b).
// Each corner deposits round(w m 2^32) (FP64 weight product, contraction
// off); the 8th corner (1, 1, 1) takes the remainder, so every particle
// deposits exactly round(m 2^32).  acc_row: the row's int64 accumulators in
// LDS (integer adds: order-independent).
template <bool FAST>
__device__ __forceinline__ void row_deposit_t(const float4 &xp, int a, int b, double mass, double scale,
                                              const double inv_h[3], int mx, int my, int mz,
                                              unsigned long long *acc_row) {
#pragma clang fp contract(off)
    const CicAxis cx = cic_axis((double)xp.x, inv_h[0], mx);
    const CicAxis cy = cic_axis((double)xp.y, inv_h[1], my);
    const CicAxis cz = cic_axis((double)xp.z, inv_h[2], mz);
    const double wxy = (a ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0);
    const long long q0 = fixed_round<FAST>(((wxy * cz.w0) * mass) * scale);
    long long q1;
    if (a & b) {
        long long s7 = 0;
#pragma unroll
        for (int e = 0; e < 7; ++e) {
            const int ea = e >> 2, eb = (e >> 1) & 1, ed = e & 1;
            const double wgt = ((ea ? cx.w1 : cx.w0) * (eb ? cy.w1 : cy.w0)) * (ed ? cz.w1 : cz.w0);
            s7 += fixed_round<FAST>((wgt * mass) * scale);
        }
        q1 = fixed_round<FAST>(mass * scale) - s7;
    } else {
        q1 = fixed_round<FAST>(((wxy * cz.w1) * mass) * scale);
    }
    atomicAdd(acc_row + cz.i0, (unsigned long long)q0);
    atomicAdd(acc_row + cz.i1, (unsigned long long)q1);
}

// Every deposit is |w m 2^32| <= |m 2^32| (CIC weights and their products lie
// in [0, 1]), so one test of the particle's mass picks the conversion.
__device__ __forceinline__ void row_deposit(const float4 &xp, int a, int b, double mass, double scale,
                                            const double inv_h[3], int mx, int my, int mz,
                                            unsigned long long *acc_row) {
    if (fixed_fast(mass * scale)) row_deposit_t<true>(xp, a, b, mass, scale, inv_h, mx, my, mz, acc_row);
    else row_deposit_t<false>(xp, a, b, mass, scale, inv_h, mx, my, mz, acc_row);
}

// ---------------------------------------------------------------------------
// k-space helpers
// ---------------------------------------------------------------------------
struct KGeom {
    int mx, my, mz, zh;
    int64_t gc;
    double kfac[3];  // 2 pi / L_d
};

__device__ __forceinline__ void kvec(const KGeom &g, int64_t kk64, double k[3], double kd[3], double &wgt) {
    const int kk = (int)kk64;  // gc < 2^31 (checked when the mesh is set)
    const int plane = g.my * g.zh;
    const int ix = kk / plane;
    const int r = kk - ix * plane;
    const int iy = r / g.zh;
    const int iz = r - iy * g.zh;
    const int nx = (2 * ix < g.mx) ? ix : ix - g.mx;
    const int ny = (2 * iy < g.my) ? iy : iy - g.my;
    k[0] = g.kfac[0] * nx;
    k[1] = g.kfac[1] * ny;
    k[2] = g.kfac[2] * iz;
    // derivative transfer: zero on the Nyquist planes of even edges
    kd[0] = (2 * ix == g.mx) ? 0.0 : k[0];
    kd[1] = (2 * iy == g.my) ? 0.0 : k[1];
    kd[2] = (2 * iz == g.mz) ? 0.0 : k[2];
    wgt = (iz == 0 || 2 * iz == g.mz) ? 1.0 : 2.0;
}

// Moments kernel instance for T types (MAXT = 2, 4 or 8), its partial row
// length 2 NP + MAXT, and its block count for a half spectrum of gc
// elements (about 8 elements per thread).
inline int moments_maxt(int T) { return T <= 2 ? 2 : (T <= 4 ? 4 : 8); }
inline int moments_row_len(int T) {
    const int mt = moments_maxt(T);
    return mt * (mt + 1) + mt;
}
inline int moments_blocks(int64_t gc) { return std::min(grid_for(gc, kBlock * 8, 4096), 1024); }

// Spectral moments of the filtered densities phi~_t = H rho_t / G (pmesh
// normalisation), summed over the full spectrum through the half spectrum
// weights: P_ts = sum_k w Re(phi~_t^* phi~_s), Q_ts = sum_k w |k|^2 Re(...),
// t <= s packed row-major, then phi~_t(0).  Every quadratic Hamiltonian's
// energy and its derivatives in chi, kappa and sigma follow from them.
// Block `blk` of `nblk` strides over the half spectrum and writes one row of
// 2 NP + MAXT doubles (the first 2 T(T+1)/2 + T are used) to `row_out`.
template <int MAXT>
__device__ __forceinline__ void moments_block(const float2 *__restrict__ spec, const KGeom &g, int T,
                                              double sigma2, double inv_g, unsigned blk, unsigned nblk,
                                              double *__restrict__ row_out) {
    constexpr int NP = MAXT * (MAXT + 1) / 2;
    double acc[2 * NP + MAXT];
#pragma unroll
    for (int j = 0; j < 2 * NP + MAXT; ++j) acc[j] = 0.0;
    const int64_t stride = (int64_t)nblk * blockDim.x;
    for (int64_t kk = (int64_t)blk * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double hn = exp(-0.5 * sigma2 * k2) * inv_g;
        double re[MAXT], im[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            const float2 v = (t < T) ? spec[(int64_t)t * g.gc + kk] : make_float2(0.f, 0.f);
            re[t] = hn * (double)v.x;
            im[t] = hn * (double)v.y;
        }
        int q = 0;
#pragma unroll
        for (int t = 0; t < MAXT; ++t)
#pragma unroll
            for (int s = t; s < MAXT; ++s, ++q) {
                const double p = wgt * (re[t] * re[s] + im[t] * im[s]);
                acc[q] += p;
                acc[NP + q] += k2 * p;
            }
        if (kk == 0)
#pragma unroll
            for (int t = 0; t < MAXT; ++t) acc[2 * NP + t] += re[t];
    }
    // compact to the caller's T and reduce
    double out[2 * NP + MAXT];
    int o = 0;
    for (int t = 0; t < T; ++t)
        for (int s = t; s < T; ++s) out[o++] = acc[t * MAXT - t * (t - 1) / 2 + (s - t)];
    for (int t = 0; t < T; ++t)
        for (int s = t; s < T; ++s) out[o++] = acc[NP + t * MAXT - t * (t - 1) / 2 + (s - t)];
    for (int t = 0; t < T; ++t) out[o++] = acc[2 * NP + t];
    for (int j = o; j < 2 * NP + MAXT; ++j) out[j] = 0.0;
    block_reduce_store<2 * NP + MAXT>(out, row_out);
}

}  // namespace hymd
