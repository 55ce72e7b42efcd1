// Particle-field path: deterministic CIC paint, fused filter / W'(phi) /
// spectral-gradient k-space pass, CIC readout, PME Poisson pass.
//
// Reference: field.py:152-200 (compute_field_force), :241-403
// (update_field_force_q), :428-616 (update_field), :692-699 (energies);
// hamiltonian.py:57-71 (filter), :145-146, :259-260, :393-412 (W).
//
// Because every Hamiltonian is quadratic, v_t = sum_s A_ts phi~_s + b_t is
// linear in the filtered densities, so the reference's real-space chain
// (filter -> c2r -> W' -> r2c -> filter -> -ik) collapses into one k-space
// pass: V_t(k) = H(k) sum_s A_ts H(k) rho_s(k)/G (+ b_t at k = 0) and
// F_{t,d}(k) = -i k_d V_t(k).  That takes T r2c + 3T c2r per update instead
// of the reference's 6T transforms (field.py:573-616).
#include <cmath>
#include <cstring>
#include <vector>

#include "fft.h"
#include "field_dev.h"

namespace hymd {

// ---------------------------------------------------------------------------
// paint
// ---------------------------------------------------------------------------
// One thread per particle; each corner deposits round(w*m*2^32) with a 64-bit
// integer atomic (order-independent => bit-reproducible grids).  The 8th
// corner takes the remainder so each particle deposits exactly round(m*2^32).
template <typename PosT>
__global__ void __launch_bounds__(kBlock) paint_fixed_kernel(const PosT *__restrict__ x,
                                                              const int32_t *__restrict__ types,
                                                              const float *__restrict__ weights, int64_t n,
                                                              Mesh m, PaintArgs pa,
                                                              unsigned long long *__restrict__ grid) {
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int t = types ? types[i] : 0;
        if ((unsigned)t >= (unsigned)pa.n_types) continue;
        const double mass = weights ? (double)weights[i] : pa.mass[t];
        const CicAxis cx = cic_axis((double)x[3 * i + 0], m.inv_h[0], m.mx);
        const CicAxis cy = cic_axis((double)x[3 * i + 1], m.inv_h[1], m.my);
        const CicAxis cz = cic_axis((double)x[3 * i + 2], m.inv_h[2], m.mz);
        unsigned long long *g = grid + (int64_t)t * m.g;
        const long long total = __double2ll_rn(mass * pa.scale);
        long long acc = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int a = c >> 2, b = (c >> 1) & 1, d = c & 1;
            const double w = ((a ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0)) * (d ? cz.w1 : cz.w0);
            long long q;
            if (c < 7) {
                q = __double2ll_rn((w * mass) * pa.scale);
                acc += q;
            } else {
                q = total - acc;
            }
            if (q != 0) {
                const int64_t idx = ((int64_t)(a ? cx.i1 : cx.i0) * m.my + (b ? cy.i1 : cy.i0)) * m.mz +
                                    (d ? cz.i1 : cz.i0);
                atomicAdd(g + idx, (unsigned long long)q);
            }
        }
    }
}

__global__ void fixed_to_float_kernel(const long long *__restrict__ in, float *__restrict__ out, int64_t n,
                                      double scale) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (float)((double)in[i] * scale);
}

__global__ void zero_kernel(unsigned long long *__restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

// ---------------------------------------------------------------------------
// Spatial binning: particles are counting-sorted by the mesh row (ix, iy) of
// their CIC base cell, so the particles whose corners can land in a row sit
// in four contiguous ranges.  Paint is then a per-row gather, readout walks
// the particles row by row.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int base_row(const float *__restrict__ x, int64_t i, const Mesh &m) {
    const CicAxis cx = cic_axis((double)x[3 * i + 0], m.inv_h[0], m.mx);
    const CicAxis cy = cic_axis((double)x[3 * i + 1], m.inv_h[1], m.my);
    return cx.i0 * m.my + cy.i0;
}

// key = mesh row of the particle's base cell; slot = its rank among that
// row's particles (wave-aggregated atomics), so the scatter needs none.
// Particles of zero weight (w != NULL) deposit nothing and are left out
// (key -1).  The loop is uniform across the wavefront for the aggregation.
__global__ void __launch_bounds__(kBlock) row_key_kernel(const float *__restrict__ x, const float *__restrict__ w,
                                                          int64_t n, Mesh m, int2 *__restrict__ keys,
                                                          int *__restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x; i0 < n; i0 += stride) {
        const int64_t i = i0 + threadIdx.x;
        int key = -1;
        if (i < n && !(w && w[i] == 0.f)) key = base_row(x, i, m);
        const int rank = rank_in_bin_all(key, counts);
        if (i < n) keys[i] = make_int2(key, rank);
    }
}

__global__ void __launch_bounds__(kBlock) key_scatter_kernel(const float *__restrict__ x,
                                                              const int32_t *__restrict__ types,
                                                              const float *__restrict__ w, int64_t n,
                                                              const int2 *__restrict__ keys,
                                                              const int *__restrict__ start, int *__restrict__ perm,
                                                              float4 *__restrict__ rec, float *__restrict__ ws) {
    // consecutive particles land in neighbouring rows: blocks are XCD-remapped
    // so that the partial cache lines of one output range fill in one L2
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int2 kr = keys[i];
        if (kr.x < 0) continue;  // left out of the binning
        const int pos = start[kr.x] + kr.y;
        perm[pos] = (int)i;
        rec[pos] = make_float4(x[3 * i], x[3 * i + 1], x[3 * i + 2], __int_as_float(types ? types[i] : 0));
        if (ws) ws[pos] = w[i];
    }
}

int launch_key_scatter(const float *x, const int32_t *types, const float *w, int64_t n, const int2 *keys,
                       const int *start, int *perm, float4 *rec, float *ws, hipStream_t s) {
    if (n > 0) key_scatter_kernel<<<grid_for(n), kBlock, 0, s>>>(x, types, w, n, keys, start, perm, rec, ws);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// One workgroup per mesh row (ix, iy) and chunk of types [t0, t0 + tc): the
// particles binned in rows (ix - a, iy - b), a, b in {0, 1}, add their corner
// deposits (a, b, d) at z = iz + d of this row into int64 LDS accumulators.
// The deposits are the fixed-point values of the scatter paint (FP64 weight
// products, contraction off; the 8th corner takes the remainder so every
// particle deposits exactly round(m 2^32)), and integer addition makes the
// sums independent of order.  The row is then written once per type as
// rho = sum * out_scale (and the int64 sums if out_fixed != NULL).
__global__ void __launch_bounds__(kBlock) paint_rows_kernel(const float4 *__restrict__ rec,
                                                             const float *__restrict__ ws,
                                                             const int *__restrict__ start, Mesh m, PaintArgs pa,
                                                             int tc, double out_scale, float *__restrict__ out,
                                                             long long *__restrict__ out_fixed) {
#pragma clang fp contract(off)
    extern __shared__ unsigned long long acc[];  // tc * mz
    const int row = (int)xcd_block(blockIdx.x, gridDim.x);
    const int t0 = (int)blockIdx.y * tc;
    const int nt = min(tc, pa.n_types - t0);
    const int ix = row / m.my, iy = row - ix * m.my;
    for (int i = threadIdx.x; i < nt * m.mz; i += blockDim.x) acc[i] = 0ull;
    int rs[4], rn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = q >> 1, b = q & 1;
        const int bx = a ? (ix == 0 ? m.mx - 1 : ix - 1) : ix;
        const int by = b ? (iy == 0 ? m.my - 1 : iy - 1) : iy;
        const int br = bx * m.my + by;
        rs[q] = start[br];
        rn[q] = start[br + 1] - rs[q];
    }
    __syncthreads();
    const int tot = rn[0] + rn[1] + rn[2] + rn[3];
    for (int v = threadIdx.x; v < tot; v += blockDim.x) {
        int q = 0, off = v;
        while (q < 3 && off >= rn[q]) {
            off -= rn[q];
            ++q;
        }
        const int p = rs[q] + off;
        const float4 r = rec[p];
        const int t = __float_as_int(r.w) - t0;
        if ((unsigned)t >= (unsigned)nt) continue;
        const int a = q >> 1, b = q & 1;
        const double mass = ws ? (double)ws[p] : pa.mass[t0 + t];
        row_deposit(r, a, b, mass, pa.scale, m.inv_h, m.mx, m.my, m.mz, acc + t * m.mz);
    }
    __syncthreads();
    const int64_t rbase = (int64_t)row * m.mz;
    for (int i = threadIdx.x; i < nt * m.mz; i += blockDim.x) {
        const int t = i / m.mz, z = i - t * m.mz;
        const int64_t o = (int64_t)(t0 + t) * m.g + rbase + z;
        const long long v = (long long)acc[i];
        out[o] = (float)((double)v * out_scale);
        if (out_fixed) out_fixed[o] = v;
    }
}

// The same row paint with one wavefront per mesh row and kPaintWaveRows rows
// per workgroup: more rows in flight per CU for rows with few particles.  A
// wavefront issues the record loads of up to four visits per lane before their
// deposits.  Same deposits, so the same integer sums and bits.
constexpr int kPaintWaveRows = kBlock / 64;

__global__ void __launch_bounds__(kBlock) paint_rows_wave_kernel(const float4 *__restrict__ rec,
                                                                  const float *__restrict__ ws,
                                                                  const int *__restrict__ start, Mesh m, PaintArgs pa,
                                                                  int tc, double out_scale, float *__restrict__ out,
                                                                  long long *__restrict__ out_fixed) {
#pragma clang fp contract(off)
    extern __shared__ unsigned long long acc_all[];  // kPaintWaveRows x tc x mz
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nrows = m.mx * m.my;
    const int row = (int)xcd_block(blockIdx.x, gridDim.x) * kPaintWaveRows + wave;
    const bool active = row < nrows;
    const int t0 = (int)blockIdx.y * tc;
    const int nt = min(tc, pa.n_types - t0);
    unsigned long long *acc = acc_all + (int64_t)wave * tc * m.mz;
    const int ix = (active ? row : 0) / m.my, iy = (active ? row : 0) - ix * m.my;
    for (int i = lane; i < nt * m.mz; i += 64) acc[i] = 0ull;
    int rs[4], rn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = q >> 1, b = q & 1;
        const int bx = a ? (ix == 0 ? m.mx - 1 : ix - 1) : ix;
        const int by = b ? (iy == 0 ? m.my - 1 : iy - 1) : iy;
        const int br = bx * m.my + by;
        rs[q] = active ? start[br] : 0;
        rn[q] = active ? start[br + 1] - rs[q] : 0;
    }
    __syncthreads();
    const int tot = rn[0] + rn[1] + rn[2] + rn[3];
    for (int v0 = 0; v0 < tot; v0 += 4 * 64) {
        float4 r[4];
        int qq[4], pp[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int v = v0 + u * 64 + lane;
            qq[u] = -1;
            if (v < tot) {
                int q = 0, off = v;
                while (q < 3 && off >= rn[q]) {
                    off -= rn[q];
                    ++q;
                }
                qq[u] = q;
                pp[u] = rs[q] + off;
                r[u] = rec[pp[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (qq[u] < 0) continue;
            const int t = __float_as_int(r[u].w) - t0;
            if ((unsigned)t >= (unsigned)nt) continue;
            const double mass = ws ? (double)ws[pp[u]] : pa.mass[t0 + t];
            row_deposit(r[u], qq[u] >> 1, qq[u] & 1, mass, pa.scale, m.inv_h, m.mx, m.my, m.mz, acc + t * m.mz);
        }
    }
    __syncthreads();
    if (!active) return;
    const int64_t rbase = (int64_t)row * m.mz;
    for (int i = lane; i < nt * m.mz; i += 64) {
        const int t = i / m.mz, z = i - t * m.mz;
        const int64_t o = (int64_t)(t0 + t) * m.g + rbase + z;
        const long long v = (long long)acc[i];
        out[o] = (float)((double)v * out_scale);
        if (out_fixed) out_fixed[o] = v;
    }
}

// ---------------------------------------------------------------------------
// readout: out[i, c] = w_i * sum_corners W * grid[type_i * ncomp + c][corner]
// ---------------------------------------------------------------------------
// perm != NULL: particle i of the (sorted) inputs is written to row perm[i].
// Blocks are XCD-remapped so that consecutive (cell-sorted) particle ranges,
// which read the same grid rows, share one L2.
// Particles come as binned records (x, y, z, type bits) or, with rec == NULL,
// as plain xyz rows of type 0.
// SORTED (binned records, no device count): each block first orders its
// particles by (type, z line of the grid) with a counting sort in LDS, so the
// lanes of a wavefront gather from few grids and few cache lines; thread k
// then reads out the block's k-th particle in that order.  Every particle's
// arithmetic and output row are those of the unsorted form.
template <int NCOMP, bool SORTED>
__global__ void __launch_bounds__(kBlock) readout_kernel(const float *__restrict__ grids,
                                                          const float4 *__restrict__ rec,
                                                          const float *__restrict__ xyz,
                                                          const float *__restrict__ scale_per_particle,
                                                          int64_t n, int n_types, Mesh m,
                                                          float *__restrict__ out, const int *__restrict__ perm,
                                                          const int *__restrict__ n_dev, int accumulate, int zshift,
                                                          int nzl) {
    __shared__ int cnt[SORTED ? kSortBuckets + 1 : 1];
    __shared__ int order[SORTED ? kBlock : 1];
    int64_t first, stride, end = n;
    if (SORTED) {  // block-uniform chunks of blockDim.x particles
        first = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
        stride = (int64_t)gridDim.x * blockDim.x;
    } else if (n_dev && gridDim.x >= kXcds) {
        // a device count of kept particles (a weighted binning): XCD x takes
        // the x-th eighth of the kept range with its own blocks (the hardware
        // deals blocks to the XCDs round robin), so each L2 sees neighbouring
        // rows and every XCD has work whatever the count
        const int64_t nk = min(n, (int64_t)*n_dev);
        const unsigned x = blockIdx.x % kXcds, xb = blockIdx.x / kXcds, nbx = gridDim.x / kXcds;
        if (xb >= nbx) return;  // the grid's remainder beyond a multiple of kXcds
        first = nk * x / kXcds + (int64_t)xb * blockDim.x + threadIdx.x;
        end = nk * (x + 1) / kXcds;
        stride = (int64_t)nbx * blockDim.x;
    } else {
        if (n_dev) end = min(n, (int64_t)*n_dev);
        first = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        stride = (int64_t)gridDim.x * blockDim.x;
    }
    for (int64_t it = first; it < end; it += stride) {
        int64_t i = it;
        if constexpr (SORTED) {
            const int64_t ip = it + threadIdx.x;
            int bk = 0;
            if (ip < n) {
                const float4 rp = rec[ip];
                const int tp = __float_as_int(rp.w);
                bk = (unsigned)tp >= (unsigned)n_types
                         ? n_types * nzl  // the last bucket: foreign types
                         : tp * nzl + min(z_line(rp.z, m.inv_h[2], m.mz, zshift), nzl - 1);
            }
            const int k = block_bucket_order(ip < n, bk, n_types * nzl + 1, cnt, order);
            if (k < 0) continue;
            i = it + k;
        }
        const float4 r = rec ? rec[i] : make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 0.f);
        const int t = __float_as_int(r.w);
        const int64_t o = perm ? (int64_t)perm[i] : i;
        if ((unsigned)t >= (unsigned)n_types) {
            if (!accumulate) {
#pragma unroll
                for (int c = 0; c < NCOMP; ++c) out[NCOMP * o + c] = 0.f;
            }
            continue;
        }
        const CicAxis cx = cic_axis((double)r.x, m.inv_h[0], m.mx);
        const CicAxis cy = cic_axis((double)r.y, m.inv_h[1], m.my);
        const CicAxis cz = cic_axis((double)r.z, m.inv_h[2], m.mz);
        const float *gt = grids + (int64_t)t * NCOMP * m.g;
        double acc[NCOMP];
#pragma unroll
        for (int q = 0; q < NCOMP; ++q) acc[q] = 0.0;
        // two halves of four corners (a = 0, then a = 1), the sums in corner
        // order; the empty asm makes the second half's x index depend on the
        // first half's sums, so its loads follow them and a thread holds 12
        // loaded values, not 24
        int ix1 = cx.i1;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float v[4][NCOMP];
            if (h == 1) {
                if constexpr (NCOMP == 3)
                    asm volatile("" : "+v"(ix1) : "v"(acc[0]), "v"(acc[1]), "v"(acc[2]));
                else
                    asm volatile("" : "+v"(ix1) : "v"(acc[0]));
            }
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4) {
                const int b = c4 >> 1, d = c4 & 1;
                const int64_t idx =
                    ((int64_t)(h ? ix1 : cx.i0) * m.my + (b ? cy.i1 : cy.i0)) * m.mz + (d ? cz.i1 : cz.i0);
#pragma unroll
                for (int q = 0; q < NCOMP; ++q) v[c4][q] = gt[q * m.g + idx];
            }
#pragma unroll
            for (int c4 = 0; c4 < 4; ++c4) {
                const int b = c4 >> 1, d = c4 & 1;
                const double w = ((h ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0)) * (d ? cz.w1 : cz.w0);
#pragma unroll
                for (int q = 0; q < NCOMP; ++q) acc[q] += w * (double)v[c4][q];
            }
        }
        const double s = scale_per_particle ? (double)scale_per_particle[i] : 1.0;
#pragma unroll
        for (int q = 0; q < NCOMP; ++q) {
            const float f = (float)(acc[q] * s);
            // accumulate: out += f, the float32 addition of a separate sum pass
            out[NCOMP * o + q] = accumulate ? out[NCOMP * o + q] + f : f;
        }
    }
}

struct FieldK {
    int T;
    int want_energy;
    double sigma2;   // sigma^2
    double inv_g;    // 1/G (r2c normalisation)
    double c0;
    double Ad[HYMD_MAX_TYPES * HYMD_MAX_TYPES];
    double b[HYMD_MAX_TYPES];
};

// Input: spec[t] = unnormalised FFT of rho_t = paint_t/dV.  Outputs:
// grad[3t+d] = -i k_d V_t(k), phif[t] = H rho_t/G (optional), spec[t] <-
// V_t(k) (optional, in place), per-block energy partial sums.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_kspace_kernel(float2 *__restrict__ spec,
                                                               float2 *__restrict__ grad,
                                                               float2 *__restrict__ phif, int write_v, KGeom g,
                                                               FieldK p, double *__restrict__ partials) {
    double e_acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const float H = (float)exp(-0.5 * p.sigma2 * k2);
        const float hn = (float)((double)H * p.inv_g);
        float2 ph[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t)
            if (t < p.T) ph[t] = cscale(spec[(int64_t)t * g.gc + kk], hn);
        if (phif) {
#pragma unroll
            for (int t = 0; t < MAXT; ++t)
                if (t < p.T) phif[(int64_t)t * g.gc + kk] = ph[t];
        }
        if (p.want_energy) {
            double e = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; ++t) {
                if (t >= p.T) continue;
#pragma unroll
                for (int s = 0; s < MAXT; ++s) {
                    if (s >= p.T) continue;
                    e += 0.5 * p.Ad[t * HYMD_MAX_TYPES + s] *
                         ((double)ph[t].x * (double)ph[s].x + (double)ph[t].y * (double)ph[s].y);
                }
            }
            e *= wgt;
            if (kk == 0) {
                for (int t = 0; t < p.T; ++t) e += p.b[t] * (double)ph[t].x;
                e += p.c0;
            }
            e_acc[0] += e;
        }
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= p.T) continue;
            float2 v = make_float2(0.f, 0.f);
#pragma unroll
            for (int s = 0; s < MAXT; ++s)
                if (s < p.T) {
                    const float a = (float)p.Ad[t * HYMD_MAX_TYPES + s];
                    v.x += a * ph[s].x;
                    v.y += a * ph[s].y;
                }
            v = cscale(v, H);
            if (kk == 0) v.x += (float)p.b[t];
            // -i k_d v = (k_d v.y, -k_d v.x)
            float2 *gt = grad + (int64_t)(3 * t) * g.gc + kk;
            gt[0] = make_float2((float)kd[0] * v.y, -(float)kd[0] * v.x);
            gt[g.gc] = make_float2((float)kd[1] * v.y, -(float)kd[1] * v.x);
            gt[2 * g.gc] = make_float2((float)kd[2] * v.y, -(float)kd[2] * v.x);
            if (write_v) spec[(int64_t)t * g.gc + kk] = v;
        }
    }
    if (partials) block_reduce_store<1>(e_acc, partials + blockIdx.x);
}

// Forward x pass fused with the k-space pass.  One workgroup owns W
// consecutive (y, kz) lines of all T density spectra (after their z and y
// passes): it runs their forward FFTs along x, then the k-space pass of
// field_kspace_kernel at every k of those lines (same arithmetic, so the same
// bits) and writes the 3T gradient spectra -i k_d V_t for the inverse passes,
// so the T transformed spectra never make a round trip through HBM.  LDS:
// twiddles, two buffers of T W lines.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_xline_kernel(const float2 *__restrict__ spec,
                                                              float2 *__restrict__ grad, KGeom g, FieldK p,
                                                              FftAxis ax, const float2 *__restrict__ tw_glob,
                                                              int wshift, double *__restrict__ partials) {
    extern __shared__ float2 smem[];
    const int n = ax.n;  // Mx
    const int W = 1 << wshift;
    const int T = p.T;
    float2 *tw = smem;
    float2 *A = tw + n;
    float2 *B = A + T * W * ax.ld;
    const int plane = g.my * g.zh;  // lines, and the element stride along x
    const int line0 = (int)xcd_block(blockIdx.x, gridDim.x) * W;
    const int wv = min(W, plane - line0);
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) tw[i] = tw_glob[i];
    // load T x W lines, W consecutive lines contiguous in memory
    const int nel = T * W * n;
    for (int i0 = threadIdx.x; i0 < nel; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = make_float2(0.f, 0.f);
            if (idx < nel) {
                const int l = idx & (W - 1), te = idx >> wshift, t = fdiv(te, ax.ndiv), e = te - t * n;
                if (l < wv) v[j] = spec[(int64_t)t * g.gc + (int64_t)e * plane + line0 + l];
            }
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < nel) {
                const int l = idx & (W - 1), te = idx >> wshift, t = fdiv(te, ax.ndiv), e = te - t * n;
                A[(t * W + l) * ax.ld + e] = v[j];
            }
        }
    }
    __syncthreads();
    const float2 *F = fft_lds(A, B, T * W, ax, tw, -1.f);
    // k-space pass, writing -i k_d V_t(k)
    double e_acc[1] = {0.0};
    for (int idx = threadIdx.x; idx < W * n; idx += blockDim.x) {
        const int l = idx & (W - 1), e = idx >> wshift;
        if (l >= wv) continue;
        const int64_t kk = (int64_t)e * plane + line0 + l;
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const float H = (float)exp(-0.5 * p.sigma2 * k2);
        const float hn = (float)((double)H * p.inv_g);
        float2 ph[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t)
            if (t < T) ph[t] = cscale(F[(t * W + l) * ax.ld + e], hn);
        if (p.want_energy) {
            double en = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; ++t) {
                if (t >= T) continue;
#pragma unroll
                for (int u = 0; u < MAXT; ++u) {
                    if (u >= T) continue;
                    en += 0.5 * p.Ad[t * HYMD_MAX_TYPES + u] *
                          ((double)ph[t].x * (double)ph[u].x + (double)ph[t].y * (double)ph[u].y);
                }
            }
            en *= wgt;
            if (kk == 0) {
#pragma unroll
                for (int t = 0; t < MAXT; ++t)
                    if (t < T) en += p.b[t] * (double)ph[t].x;
                en += p.c0;
            }
            e_acc[0] += en;
        }
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= T) continue;
            float2 v = make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
                if (u < T) {
                    const float a = (float)p.Ad[t * HYMD_MAX_TYPES + u];
                    v.x += a * ph[u].x;
                    v.y += a * ph[u].y;
                }
            v = cscale(v, H);
            if (kk == 0) v.x += (float)p.b[t];
            float2 *gt = grad + (int64_t)(3 * t) * g.gc + kk;
            gt[0] = make_float2((float)kd[0] * v.y, -(float)kd[0] * v.x);
            gt[g.gc] = make_float2((float)kd[1] * v.y, -(float)kd[1] * v.x);
            gt[2 * g.gc] = make_float2((float)kd[2] * v.y, -(float)kd[2] * v.x);
        }
    }
    if (partials) block_reduce_store<1>(e_acc, partials + blockIdx.x);
}

// Forward x pass, k-space pass and the inverse x passes in one workgroup.
// k_y and k_z are constant along an x line, so the x-inverse of -i k_d V_t
// for d = y, z is -i k_d times the x-inverse of V_t: per line and type only
// two inverse transforms are needed, IFFT_x(-i k_x V_t) -> out[t] and
// IFFT_x(V_t) -> out[T + t]; fft_c2r_yz_grad applies -i k_y / -i k_z in its
// y-z plane pass.  The T spectra are transformed in place in one LDS buffer
// of T W lines; V_t waits in registers (PMAX positions per thread) while the
// x-gradient family is transformed and stored.
// x-line geometry: runtime (Mx, W, T from the plan) or compile-time (Mx = N,
// T types, W lines per workgroup), the latter with every index map a constant
// multiply-shift and the Stockham stages of fft_inplace_ct.
template <int ELEMS>
struct XlineRt {
    static constexpr bool kCt = false;
    int n, W, wshift, T, ld;
    FastDiv ndiv, flines;
    FftAxis ax;
    __device__ int div_n(int i) const { return fdiv(i, ndiv); }
    template <int SGN, int NT>
    __device__ void fft(float2 *D, const float2 *tw) const {
        fft_inplace<ELEMS>(D, 1, ld, T * W, flines, ax, tw, (float)SGN);
    }
};

template <int T_, int N, int W_>
struct XlineCt {
    static constexpr int n = N, W = W_, T = T_, ld = N + 1;
    static constexpr int wshift = W_ == 1 ? 0 : W_ == 2 ? 1 : W_ == 4 ? 2 : W_ == 8 ? 3 : 4;
    static constexpr bool kCt = true;
    __device__ int div_n(int i) const { return i / N; }
    template <int SGN, int NT>
    __device__ void fft(float2 *D, const float2 *tw) const {
        fft_inplace_ct<N, T_ * W_, 1, N + 1, NT, SGN>(D, tw);
    }
    // the same with the first stage's inputs from ld(line, e) / the last
    // stage's outputs to st(line, e, value), line = t W + l
    template <int SGN, int NT, class Load>
    __device__ void fft_from(float2 *D, const float2 *tw, Load ld) const {
        fft_first_ct<N, T_ * W_, 1, N + 1, NT, SGN>(D, ld);
        fft_inplace_ct<N, T_ * W_, 1, N + 1, NT, SGN, 1>(D, tw);
    }
    template <int SGN, int NT, class Store>
    __device__ void fft_to(float2 *D, const float2 *tw, Store st) const {
        fft_inplace_ct<N, T_ * W_, 1, N + 1, NT, SGN, 0, CtPlan<N>::v.ns - 1>(D, tw);
        fft_last_ct<N, T_ * W_, 1, N + 1, NT, SGN, false>(D, tw, st);
    }
};

struct PmeK {
    double sigma2;
    double inv_g;
    double four_pi_elec;  // 4 pi k_e / eps_r
};

// PME (T = 1): the k-space pass is pme_kspace_kernel's (psi = 4 pi elec H
// rho~ / |k|^2 with the same FP64 arithmetic); outputs -i k_x psi and psi.
template <int MAXT, int NT, class G, bool PME = false>
__global__ void __launch_bounds__(NT) field_xline_grad_kernel(const float2 *__restrict__ spec,
                                                              float2 *__restrict__ out, KGeom g, FieldK p, G geo,
                                                              const float2 *__restrict__ tw_glob,
                                                              double *__restrict__ partials, PmeK pk) {
    constexpr int PMAX = 16 / MAXT;  // W n <= NT PMAX (checked on the host)
    extern __shared__ float2 smem[];
    const int n = geo.n;  // Mx
    const int W = geo.W, wshift = geo.wshift;
    const int T = geo.T;
    const int ld = geo.ld;
    float2 *tw = smem;
    float2 *D = tw + n;  // line (t W + l) at D + (t W + l) ld
    const int plane = g.my * g.zh;
    const int line0 = (int)xcd_block(blockIdx.x, gridDim.x) * W;
    const int wv = min(W, plane - line0);
    for (int i = threadIdx.x; i < n - 1; i += NT) tw[i] = tw_glob[i];
    const int nel = T * W * n;
    if constexpr (G::kCt) {
        // the forward transform's first stage reads the lines from global memory
        geo.template fft_from<-1, NT>(D, tw, [&](int line, int e) -> float2 {
            const int t = line >> wshift, l = line & (W - 1);
            return l < wv ? spec[(int64_t)t * g.gc + (int64_t)e * plane + line0 + l] : make_float2(0.f, 0.f);
        });
    } else {
    for (int i0 = threadIdx.x; i0 < nel; i0 += NT * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * NT;
            v[j] = make_float2(0.f, 0.f);
            if (idx < nel) {
                const int l = idx & (W - 1), te = idx >> wshift, t = geo.div_n(te), e = te - t * n;
                if (l < wv) v[j] = spec[(int64_t)t * g.gc + (int64_t)e * plane + line0 + l];
            }
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * NT;
            if (idx < nel) {
                const int l = idx & (W - 1), te = idx >> wshift, t = geo.div_n(te), e = te - t * n;
                D[(t * W + l) * ld + e] = v[j];
            }
        }
    }
    __syncthreads();
    geo.template fft<-1, NT>(D, tw);
    }
    // k-space pass (the arithmetic of field_kspace_kernel): V_t kept in
    // registers, -i k_x V_t written back in place
    double e_acc[1] = {0.0};
    float2 hold[PMAX][MAXT];
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        const int pos = threadIdx.x + j * NT;
        const int l = pos & (W - 1), e = pos >> wshift;
        if (pos >= W * n || l >= wv) {
#pragma unroll
            for (int t = 0; t < MAXT; ++t) hold[j][t] = make_float2(0.f, 0.f);
            continue;
        }
        const int64_t kk = (int64_t)e * plane + line0 + l;
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        if constexpr (PME) {
            double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
            const double H = exp(-0.5 * pk.sigma2 * k2);
            if (kk == 0) k2 = 1.0;
            const float2 sp = D[l * ld + e];
            const double rx = (double)sp.x * pk.inv_g, ry = (double)sp.y * pk.inv_g;
            const double f = pk.four_pi_elec * H / k2;
            const double px = f * rx, py = f * ry;
            e_acc[0] += 0.5 * wgt * (rx * px + ry * py);
            hold[j][0] = make_float2((float)px, (float)py);
            D[l * ld + e] = make_float2((float)(kd[0] * py), (float)(-kd[0] * px));
            continue;
        }
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const float H = (float)exp(-0.5 * p.sigma2 * k2);
        const float hn = (float)((double)H * p.inv_g);
        float2 ph[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) ph[t] = t < T ? cscale(D[(t * W + l) * ld + e], hn) : make_float2(0.f, 0.f);
        if (p.want_energy) {
            double en = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; ++t) {
                if (t >= T) continue;
#pragma unroll
                for (int u = 0; u < MAXT; ++u) {
                    if (u >= T) continue;
                    en += 0.5 * p.Ad[t * HYMD_MAX_TYPES + u] *
                          ((double)ph[t].x * (double)ph[u].x + (double)ph[t].y * (double)ph[u].y);
                }
            }
            en *= wgt;
            if (kk == 0) {
#pragma unroll
                for (int t = 0; t < MAXT; ++t)
                    if (t < T) en += p.b[t] * (double)ph[t].x;
                en += p.c0;
            }
            e_acc[0] += en;
        }
        const float kx = (float)kd[0];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            float2 v = make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
                if (u < T) {
                    const float a = (float)p.Ad[t * HYMD_MAX_TYPES + u];
                    v.x += a * ph[u].x;
                    v.y += a * ph[u].y;
                }
            v = cscale(v, H);
            if (kk == 0) v.x += (float)p.b[t];
            hold[j][t] = v;
            if (t < T) D[(t * W + l) * ld + e] = make_float2(kx * v.y, -kx * v.x);
        }
    }
    __syncthreads();
    if constexpr (G::kCt) {
        // the inverse transforms' last stage writes the lines to global memory
        geo.template fft_to<1, NT>(D, tw, [&](int line, int e, float2 v) {
            const int t = line >> wshift, l = line & (W - 1);
            if (l < wv) out[(int64_t)t * g.gc + (int64_t)e * plane + line0 + l] = v;
        });
    } else {
        geo.template fft<1, NT>(D, tw);
        for (int idx = threadIdx.x; idx < nel; idx += NT) {
            const int l = idx & (W - 1), te = idx >> wshift, t = geo.div_n(te), e = te - t * n;
            if (l < wv) out[(int64_t)t * g.gc + (int64_t)e * plane + line0 + l] = D[(t * W + l) * ld + e];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PMAX; ++j) {
        const int pos = threadIdx.x + j * NT;
        if (pos >= W * n) continue;
        const int l = pos & (W - 1), e = pos >> wshift;
#pragma unroll
        for (int t = 0; t < MAXT; ++t)
            if (t < T) D[(t * W + l) * ld + e] = hold[j][t];
    }
    __syncthreads();
    if constexpr (G::kCt) {
        geo.template fft_to<1, NT>(D, tw, [&](int line, int e, float2 v) {
            const int t = line >> wshift, l = line & (W - 1);
            if (l < wv) out[(int64_t)(T + t) * g.gc + (int64_t)e * plane + line0 + l] = v;
        });
    } else {
        geo.template fft<1, NT>(D, tw);
        for (int idx = threadIdx.x; idx < nel; idx += NT) {
            const int l = idx & (W - 1), te = idx >> wshift, t = geo.div_n(te), e = te - t * n;
            if (l < wv) out[(int64_t)(T + t) * g.gc + (int64_t)e * plane + line0 + l] = D[(t * W + l) * ld + e];
        }
    }
    if (partials) block_reduce_store<1, NT>(e_acc, partials + blockIdx.x);
}

// Spectral moments (P, Q, phi~(0)) of the filtered densities, one partial
// row per block.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_moments_kernel(const float2 *__restrict__ spec, KGeom g, int T,
                                                                double sigma2, double inv_g,
                                                                double *__restrict__ partials) {
    moments_block<MAXT>(spec, g, T, sigma2, inv_g, blockIdx.x, gridDim.x,
                        partials + (int64_t)blockIdx.x * (2 * (MAXT * (MAXT + 1) / 2) + MAXT));
}

// ---------------------------------------------------------------------------
// Adjoint of the force chain F = readout(c2r(-i k_d H A H rho^)) with respect
// to the coupling A and sigma.  With the upstream gradient g = dL/dF painted
// per type t and component d (CIC paint is the adjoint of the CIC readout)
// into P_td, and D_t = sum_d i k_d FFT(P_td):
//   dL/dA_ts   = 1/G sum_k H(k)^2 Re(conj(D_t) rho^_s)           (raw FFTs)
//   dL/dsigma  = -2 sigma 1/G sum_k |k|^2 H^2 Re(conj(D_t) (A rho^)_t)
// summed over the full spectrum through the half-spectrum weights; the
// k = 0 shift b has no gradient there, so it carries no force term.
// ---------------------------------------------------------------------------

// Binned slot j of the unweighted binning: ws[d * n + j] = g[perm[j], d];
// the largest |g| goes to *maxbits as float bits (+inf for a non-finite g).
__global__ void __launch_bounds__(kBlock) adjoint_gather_kernel(const float *__restrict__ g,
                                                                 const int *__restrict__ perm, int64_t n,
                                                                 float *__restrict__ ws, unsigned *maxbits) {
    float mx = 0.f;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
        const int64_t i = perm[j];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float v = g[3 * i + d];
            ws[d * n + j] = v;
            mx = isfinite(v) ? fmaxf(mx, fabsf(v)) : INFINITY;
        }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_down(mx, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(maxbits, __float_as_uint(mx));
}

struct AdjK {
    double A[8][8];
};

// One partial row per block: [MAXT x MAXT] sum w H^2 Re(conj(D_t) rho_s), then
// sum w |k|^2 H^2 Re(conj(D_t) (A rho)_t).  spec: T raw density spectra;
// pspec: 3T raw spectra of the painted gradient, component-major [d][t].
template <int MAXT>
__global__ void __launch_bounds__(kBlock) adjoint_moments_kernel(const float2 *__restrict__ spec,
                                                                  const float2 *__restrict__ pspec, KGeom g, int T,
                                                                  double sigma2, AdjK p,
                                                                  double *__restrict__ partials) {
    constexpr int NV = MAXT * MAXT + 1;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double h2 = wgt * exp(-sigma2 * k2);
        double rr[MAXT], ri[MAXT];
#pragma unroll
        for (int s = 0; s < MAXT; ++s) {
            const float2 v = (s < T) ? spec[(int64_t)s * g.gc + kk] : make_float2(0.f, 0.f);
            rr[s] = v.x;
            ri[s] = v.y;
        }
        double sig = 0.0;
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= T) continue;
            // D_t = sum_d i k_d P_td
            double dr = 0.0, di = 0.0;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float2 q = pspec[((int64_t)d * T + t) * g.gc + kk];
                dr -= kd[d] * (double)q.y;
                di += kd[d] * (double)q.x;
            }
            double arr = 0.0, ari = 0.0;
#pragma unroll
            for (int s = 0; s < MAXT; ++s) {
                acc[t * MAXT + s] += h2 * (dr * rr[s] + di * ri[s]);
                arr += p.A[t][s] * rr[s];
                ari += p.A[t][s] * ri[s];
            }
            sig += dr * arr + di * ari;
        }
        acc[MAXT * MAXT] += k2 * h2 * sig;
    }
    block_reduce_store<NV>(acc, partials + (int64_t)blockIdx.x * NV);
}

// comp_pressure's field terms (pressure.py:9-179) from the filtered spectra
// phi~_t (pmesh normalisation) by Parseval, summed over the full spectrum
// through the half-spectrum weights, with V_t = (A phi~)_t:
//   [0]      S   = sum_k Re(phi~^* . V)
//   [1..3]   S_d = sum_k k_d^2 Re(phi~^* . V)
//   [4]      E1  = sum_k Re(psi^* sum_t q_t phi~_t)            (psi != NULL)
//   [5..7]   E1_d = sum_k k_d^2 Re(psi^* sum_t q_t phi~_t)
//   [8]      Eq  = sum_k Re(phi_q^* psi)                       (phi_q != NULL)
//   [9..9+T) phi~_t(k = 0)
// k_d is -pi/h on the Nyquist plane of an even edge (its square is what the
// Laplacian transfer uses).  psi and phi_q are raw r2c spectra, scaled by
// inv_g here.
constexpr int kPressNv = 9 + 8;
struct PressK {
    double A[8][8];
    double q[8];
    double inv_g;
};

template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_pressure_kernel(const float2 *__restrict__ phif,
                                                                 const float2 *__restrict__ psif,
                                                                 const float2 *__restrict__ qf, KGeom g, int T,
                                                                 PressK p, double *__restrict__ partials) {
    double acc[kPressNv];
#pragma unroll
    for (int j = 0; j < kPressNv; ++j) acc[j] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        double re[MAXT], im[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            const float2 v = (t < T) ? phif[(int64_t)t * g.gc + kk] : make_float2(0.f, 0.f);
            re[t] = (double)v.x;
            im[t] = (double)v.y;
        }
        double sw = 0.0;
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            double vr = 0.0, vi = 0.0;
#pragma unroll
            for (int u = 0; u < MAXT; ++u) {
                vr += p.A[t][u] * re[u];
                vi += p.A[t][u] * im[u];
            }
            sw += re[t] * vr + im[t] * vi;
        }
        const double k2[3] = {k[0] * k[0], k[1] * k[1], k[2] * k[2]};
        acc[0] += wgt * sw;
#pragma unroll
        for (int d = 0; d < 3; ++d) acc[1 + d] += wgt * k2[d] * sw;
        if (psif) {
            const float2 ps = psif[kk];
            const double pr = p.inv_g * (double)ps.x, pi = p.inv_g * (double)ps.y;
            double cr = 0.0, ci = 0.0;
#pragma unroll
            for (int t = 0; t < MAXT; ++t) {
                cr += p.q[t] * re[t];
                ci += p.q[t] * im[t];
            }
            const double e1 = pr * cr + pi * ci;
            acc[4] += wgt * e1;
#pragma unroll
            for (int d = 0; d < 3; ++d) acc[5 + d] += wgt * k2[d] * e1;
            if (qf) {
                const float2 qv = qf[kk];
                acc[8] += wgt * (p.inv_g * (double)qv.x * pr + p.inv_g * (double)qv.y * pi);
            }
        }
        if (kk == 0)
#pragma unroll
            for (int t = 0; t < MAXT; ++t) acc[9 + t] += re[t];
    }
    block_reduce_store<kPressNv>(acc, partials + (int64_t)blockIdx.x * kPressNv);
}


// PME (field.py:363-397): rho~ = H rho/G; psi = 4 pi elec rho~/|k|^2 (|k|^2 := 1
// at k = 0); E_d = -i k_d psi.  Energy partials: 1/2 wgt Re(rho^* psi).
// grad_form: out[0] = -i k_x psi, out[1] = psi (the inputs of the x-inverse
// and fft_c2r_yz_grad, which applies -i k_y and -i k_z); otherwise out[d] =
// E_d and, with write_psi, out[3] = psi.
__global__ void __launch_bounds__(kBlock) pme_kspace_kernel(const float2 *__restrict__ spec,
                                                             float2 *__restrict__ out, int write_psi,
                                                             int grad_form, KGeom g, PmeK p,
                                                             double *__restrict__ partials) {
    double e_acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const double H = exp(-0.5 * p.sigma2 * k2);
        if (kk == 0) k2 = 1.0;
        const float2 s = spec[kk];
        const double rx = (double)s.x * p.inv_g, ry = (double)s.y * p.inv_g;  // rho (normalised)
        const double f = p.four_pi_elec * H / k2;
        const double px = f * rx, py = f * ry;                                // psi
        e_acc[0] += 0.5 * wgt * (rx * px + ry * py);
        out[kk] = make_float2((float)(kd[0] * py), (float)(-kd[0] * px));
        if (grad_form) {
            out[g.gc + kk] = make_float2((float)px, (float)py);
            continue;
        }
        out[g.gc + kk] = make_float2((float)(kd[1] * py), (float)(-kd[1] * px));
        out[2 * g.gc + kk] = make_float2((float)(kd[2] * py), (float)(-kd[2] * px));
        if (write_psi) out[3 * g.gc + kk] = make_float2((float)px, (float)py);
    }
    if (partials) block_reduce_store<1>(e_acc, partials + blockIdx.x);
}

// dV sum w(phi~) over real cells (field.py:692-693), FP64 accumulation
template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_energy_kernel(const float *__restrict__ phi, int64_t g, FieldK p,
                                                               double *__restrict__ partials) {
    double acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g; i += stride) {
        double ph[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) ph[t] = (t < p.T) ? (double)phi[(int64_t)t * g + i] : 0.0;
        double w = p.c0;
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= p.T) continue;
            w += p.b[t] * ph[t];
#pragma unroll
            for (int s = 0; s < MAXT; ++s)
                if (s < p.T) w += 0.5 * p.Ad[t * HYMD_MAX_TYPES + s] * ph[t] * ph[s];
        }
        acc[0] += w;
    }
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

// sum (1/2 phi_q psi) over real cells (field.py:697-699)
__global__ void __launch_bounds__(kBlock) field_q_energy_kernel(const float *__restrict__ phi_q,
                                                                 const float *__restrict__ psi, int64_t g,
                                                                 double *__restrict__ partials) {
    double acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g; i += stride)
        acc[0] += 0.5 * (double)phi_q[i] * (double)psi[i];
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

__global__ void self_energy_kernel(const float *__restrict__ q, int64_t n, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc[0] += (double)q[i] * (double)q[i];
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

}  // namespace hymd

// ===========================================================================
// Handle and C-ABI
// ===========================================================================
using namespace hymd;

struct hymd_pm {
    Mesh mesh;
    int max_types;
    hipStream_t stream;
    FftPlan fft;
    DevBuf paint;     // max_types x G: float densities, or int64 for the accumulate-mode paint
    DevBuf spec;      // max_types x Gc complex
    DevBuf grad;      // max(3 max_types, 4) x Gc complex
    DevBuf cscratch;  // max_types x Gc complex (phi / v_ext inverse transforms)
    DevBuf rscratch;  // 3 x G float (PME field when the caller passes none)
    DevBuf partials;  // energy partial sums
    DevBuf result;    // reduced doubles
    HostBuf host;     // pinned readback
    int kblocks;      // blocks of the k-space passes
    DevBuf mpart;     // moments partial rows
    // spatial sort of the last painted particle set (row-binned order)
    DevBuf keys, perm, xs, ws;  // per particle (xs: binned records x, y, z, type bits)
    DevBuf counts, start, tiles;  // per mesh row (counts kept zero between sorts)
    int64_t sorted_n = 0;
    bool sorted_weights = false;
    // spec holds the unnormalised spectra of the densities of the binned set
    // (left by hymd_field_moments; any new binning clears it)
    bool density_spectra = false;
    int density_types = 0;
    DevBuf adj;       // adjoint: 3 x n binned upstream gradient components, then their max |g| bits
    // launch shapes chosen by hymd_pm_set_tuning (an autotuner's choices)
    int paint_rows = 1;       // mesh rows per paint workgroup: 1 or 4 (one per wavefront)
    int readout_block = kBlock;
    int xline_w = 0;          // lines per x-line workgroup, 0: the widest that fits
    uint64_t box_version = 0;  // hymd_pm_set_box and hymd_pm_set_tuning calls
};

uint64_t hymd::pm_generation(const hymd_pm *pm) {
    uint64_t g = pm->box_version;
    for (const DevBuf *b : {&pm->paint, &pm->spec, &pm->grad, &pm->cscratch, &pm->rscratch, &pm->partials,
                            &pm->result, &pm->mpart, &pm->keys, &pm->perm, &pm->xs, &pm->ws, &pm->counts,
                            &pm->start, &pm->tiles})
        g += b->allocs;
    return g;
}

int hymd::pm_clone(const hymd_pm *src, int32_t max_types, void *stream, hymd_pm **out) {
    const int32_t mesh[3] = {src->mesh.mx, src->mesh.my, src->mesh.mz};
    return hymd_pm_create(out, mesh, src->mesh.box, max_types, stream);
}

void hymd::pm_box(const hymd_pm *pm, double box[3]) {
    for (int d = 0; d < 3; ++d) box[d] = pm->mesh.box[d];
}

// LDS budget of one x-line workgroup (two per CU)
constexpr size_t kXlineLdsBytes = 80 * 1024;

static size_t xline_lds(const FftPlan &f, int T, int W) {
    return sizeof(float2) * ((size_t)f.ax.n + 2 * (size_t)T * W * f.ax.ld);
}

// Lines per x-line workgroup for T types (0: the fused pass does not fit).
static int xline_width(const FftPlan &f, int T) {
    for (int W = 16; W >= 2; W >>= 1)
        if (xline_lds(f, T, W) <= kXlineLdsBytes) return W;
    return 0;
}

template <int MAXT>
static int launch_xline(hymd_pm *pm, const float2 *spec, float2 *grad, const KGeom &g, const FieldK &p, int W,
                        double *partials, int *nblocks) {
    const FftPlan &f = pm->fft;
    const int plane = g.my * g.zh;
    const int nb = (plane + W - 1) / W;
    const size_t lds = xline_lds(f, p.T, W);
    static bool attr_set = false;  // the opt-in is per kernel and process
    if (!attr_set && lds > 64 * 1024) {
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&field_xline_kernel<MAXT>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kXlineLdsBytes));
        attr_set = true;
    }
    int wshift = 0;
    while ((1 << wshift) < W) ++wshift;
    field_xline_kernel<MAXT><<<nb, kBlock, lds, pm->stream>>>(spec, grad, g, p, f.ax, (const float2 *)f.tw_x.ptr,
                                                              wshift, partials);
    HYMD_HIP(hipGetLastError());
    *nblocks = nb;
    return HYMD_OK;
}

constexpr int kXlineGradThreads = 256;
constexpr int kXlineGradElems = 16;  // values per thread and in-place stage
// Workgroups of 512 or 1024 threads (one per CU at up to 150 KB of LDS) for
// long x lines with many types, where 256 threads hold only 2 lines: a
// workgroup's lines are neighbouring (y, k_z) columns, so W lines give W * 8
// contiguous bytes per x row, and at 16 bytes the partial HBM bursts double
// the traffic of the pass.
constexpr size_t kXlineWideLdsBytes = 150 * 1024;

struct XlineShape {
    int W = 0, NT = kXlineGradThreads;
};

// Lines per workgroup W of field_xline_grad_kernel with NT threads for T
// types: the registers hold PMAX positions per thread (W n <= NT PMAX), the
// in-place stages kXlineGradElems values per thread, LDS the T W lines.
static bool xline_grad_fits(const FftPlan &f, int T, int W, int NT) {
    const int pmax = 16 / moments_maxt(T);
    if ((int64_t)W * f.ax.n > (int64_t)NT * pmax) return false;
    const size_t lim = NT == kXlineGradThreads ? kXlineLdsBytes : kXlineWideLdsBytes;
    if (sizeof(float2) * ((size_t)f.ax.n + (size_t)T * W * f.ax.ld) > lim) return false;
    return fft_inplace_fits(f.ax, T * W, NT, kXlineGradElems);
}

// The x-line shape: `want` lines with the fewest threads that fit it, else
// the widest W that fits 256 threads when that is at least 4 lines, else 8
// or 16 lines with 512 or 1024 threads, else 2 lines (W = 0: the fused pass
// does not fit).
static XlineShape xline_grad_shape(const FftPlan &f, int T, int want = 0) {
    const int nts[3] = {kXlineGradThreads, 2 * kXlineGradThreads, 4 * kXlineGradThreads};
    if (want > 0)
        for (int NT : nts)
            if (xline_grad_fits(f, T, want, NT)) return XlineShape{want, NT};
    for (int W = 16; W >= 4; W >>= 1)
        if (xline_grad_fits(f, T, W, kXlineGradThreads)) return XlineShape{W, kXlineGradThreads};
    for (int W = 16; W >= 8; W >>= 1)
        for (int NT : {2 * kXlineGradThreads, 4 * kXlineGradThreads})
            if (xline_grad_fits(f, T, W, NT)) return XlineShape{W, NT};
    if (xline_grad_fits(f, T, 2, kXlineGradThreads)) return XlineShape{2, kXlineGradThreads};
    return XlineShape{};
}

template <int MAXT, int NT, class G>
static int launch_xline_grad_g(hymd_pm *pm, const float2 *spec, float2 *out, const KGeom &g, const FieldK &p, int W,
                               const G &geo, double *partials, int *nblocks, const PmeK *pme) {
    const FftPlan &f = pm->fft;
    const int plane = g.my * g.zh;
    const int nb = (plane + W - 1) / W;
    const size_t lds = sizeof(float2) * ((size_t)f.ax.n + (size_t)p.T * W * f.ax.ld);
    const size_t lim = NT == kXlineGradThreads ? kXlineLdsBytes : kXlineWideLdsBytes;
    static bool attr_set = false;  // the opt-in is per kernel and process
    if (!attr_set && lds > 64 * 1024) {
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&field_xline_grad_kernel<MAXT, NT, G>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&field_xline_grad_kernel<MAXT, NT, G, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lim));
        attr_set = true;
    }
    if (pme)
        field_xline_grad_kernel<MAXT, NT, G, true><<<nb, NT, lds, pm->stream>>>(
            spec, out, g, p, geo, (const float2 *)f.tw_x.ptr, partials, *pme);
    else
        field_xline_grad_kernel<MAXT, NT, G><<<nb, NT, lds, pm->stream>>>(spec, out, g, p, geo,
                                                                          (const float2 *)f.tw_x.ptr, partials, PmeK{});
    HYMD_HIP(hipGetLastError());
    *nblocks = nb;
    return HYMD_OK;
}

template <int MAXT, class G>
static int launch_xline_grad_nt(hymd_pm *pm, const float2 *spec, float2 *out, const KGeom &g, const FieldK &p,
                                const XlineShape &sh, const G &geo, double *partials, int *nblocks,
                                const PmeK *pme) {
    if (sh.NT == 4 * kXlineGradThreads)
        return launch_xline_grad_g<MAXT, 4 * kXlineGradThreads>(pm, spec, out, g, p, sh.W, geo, partials, nblocks,
                                                                pme);
    if (sh.NT == 2 * kXlineGradThreads)
        return launch_xline_grad_g<MAXT, 2 * kXlineGradThreads>(pm, spec, out, g, p, sh.W, geo, partials, nblocks,
                                                                pme);
    return launch_xline_grad_g<MAXT, kXlineGradThreads>(pm, spec, out, g, p, sh.W, geo, partials, nblocks, pme);
}

template <int N>
static bool xline_ct_plan(const FftPlan &f) {
    const StagePlan &sp = CtPlan<N>::v;
    if (f.ax.n != N || f.ax.ld != N + 1 || f.ax.nstages != sp.ns) return false;
    for (int i = 0; i < sp.ns; ++i)
        if (f.ax.radix[i] != sp.r[i] || f.ax.twoff[i] != sp.twoff[i]) return false;
    return true;
}

// The x-line pass for T types in the shape sh: a compile-time geometry for
// Mx = 128 or 160 at the 256-thread W that xline_grad_shape picks there (run
// at 512 threads for Mx = 128) and for Mx = 256 with six types at 8 lines of
// 1024 threads or 4 of 512, the runtime one otherwise.
static int launch_xline_grad(hymd_pm *pm, const float2 *spec, float2 *out, const KGeom &g, const FieldK &p,
                             const XlineShape &sh, double *partials, int *nblocks, const PmeK *pme = nullptr) {
    const FftPlan &f = pm->fft;
    const int T = p.T, W = sh.W;
    // NTH threads per workgroup for the 256-thread shape's lines and LDS: the
    // Mx = 128 geometries run at 512, twice the waves in flight (the 128^3,
    // T = 4 pass 33.6 -> 29.8 us); at 160 that made the peptide step slower
#define HYMD_XLINE_CT(NX, TT, WW, MT, NTH)                                                                 \
    if (sh.NT == kXlineGradThreads && f.ax.n == NX && T == TT && W == WW && xline_ct_plan<NX>(f))          \
        return launch_xline_grad_g<MT, NTH>(pm, spec, out, g, p, W, XlineCt<TT, NX, WW>{}, partials, nblocks, pme);
    HYMD_XLINE_CT(128, 1, 16, 2, 2 * kXlineGradThreads)
    HYMD_XLINE_CT(128, 2, 16, 2, 2 * kXlineGradThreads)
    HYMD_XLINE_CT(128, 3, 8, 4, 2 * kXlineGradThreads)
    HYMD_XLINE_CT(128, 4, 8, 4, 2 * kXlineGradThreads)
    HYMD_XLINE_CT(160, 1, 8, 2, kXlineGradThreads)
    HYMD_XLINE_CT(160, 2, 8, 2, kXlineGradThreads)
    HYMD_XLINE_CT(160, 3, 4, 4, kXlineGradThreads)
    HYMD_XLINE_CT(160, 4, 4, 4, kXlineGradThreads)
    HYMD_XLINE_CT(256, 1, 8, 2, kXlineGradThreads)
#undef HYMD_XLINE_CT
    // 256-point lines with six types (the 256^3 workloads): 8 lines of 1024
    // threads or 4 lines of 512
    if (f.ax.n == 256 && T == 6 && xline_ct_plan<256>(f)) {
        if (sh.NT == 4 * kXlineGradThreads && W == 8)
            return launch_xline_grad_g<8, 4 * kXlineGradThreads>(pm, spec, out, g, p, W, XlineCt<6, 256, 8>{},
                                                                 partials, nblocks, pme);
        if (sh.NT == 2 * kXlineGradThreads && W == 4)
            return launch_xline_grad_g<8, 2 * kXlineGradThreads>(pm, spec, out, g, p, W, XlineCt<6, 256, 4>{},
                                                                 partials, nblocks, pme);
    }
    int wshift = 0;
    while ((1 << wshift) < W) ++wshift;
    const XlineRt<kXlineGradElems> geo{f.ax.n, W, wshift, T, f.ax.ld, f.ax.ndiv, make_fastdiv((uint32_t)(T * W)), f.ax};
    if (T <= 2) return launch_xline_grad_nt<2>(pm, spec, out, g, p, sh, geo, partials, nblocks, pme);
    if (T <= 4) return launch_xline_grad_nt<4>(pm, spec, out, g, p, sh, geo, partials, nblocks, pme);
    return launch_xline_grad_nt<8>(pm, spec, out, g, p, sh, geo, partials, nblocks, pme);
}

// LDS budget of one row-paint workgroup (int64 accumulators of tc types x Mz)
constexpr size_t kRowLdsBytes = 32 * 1024;

// Edges <= 1024 keep G and Gc below 2^31, so kernels index cells in 32 bits.
static int set_mesh(hymd_pm *pm, const int32_t mesh[3], const double box[3]) {
    for (int d = 0; d < 3; ++d) {
        HYMD_REQUIRE(mesh[d] >= 4 && mesh[d] <= 1024 && mesh[d] % 2 == 0,
                     "mesh edge %d must be even and in [4, 1024] (got %d)", d, mesh[d]);
        HYMD_REQUIRE(box[d] > 0.0 && std::isfinite(box[d]), "box edge %d must be positive (got %g)", d, box[d]);
    }
    Mesh &m = pm->mesh;
    m.mx = mesh[0];
    m.my = mesh[1];
    m.mz = mesh[2];
    m.zh = m.mz / 2 + 1;
    m.g = (int64_t)m.mx * m.my * m.mz;
    m.gc = (int64_t)m.mx * m.my * m.zh;
    for (int d = 0; d < 3; ++d) {
        m.box[d] = box[d];
        m.inv_h[d] = (double)mesh[d] / box[d];
    }
    return HYMD_OK;
}

static KGeom kgeom(const hymd_pm *pm) {
    const Mesh &m = pm->mesh;
    KGeom g;
    g.mx = m.mx;
    g.my = m.my;
    g.mz = m.mz;
    g.zh = m.zh;
    g.gc = m.gc;
    for (int d = 0; d < 3; ++d) g.kfac[d] = 2.0 * M_PI / m.box[d];
    return g;
}

static double cell_volume(const hymd_pm *pm) {
    const Mesh &m = pm->mesh;
    return (m.box[0] * m.box[1] * m.box[2]) / (double)m.g;
}

extern "C" int hymd_pm_create(hymd_pm **out, const int32_t mesh[3], const double box[3], int32_t max_types,
                              void *stream) {
    HYMD_REQUIRE(out && mesh && box, "null argument");
    HYMD_REQUIRE(max_types >= 1 && max_types <= HYMD_MAX_TYPES, "max_types must be in [1, %d]", HYMD_MAX_TYPES);
    hymd_pm *pm = new hymd_pm();
    int st = set_mesh(pm, mesh, box);
    if (st == HYMD_OK) st = fft_plan_init(pm->fft, pm->mesh);
    pm->max_types = max_types;
    pm->stream = (hipStream_t)stream;
    const Mesh &m = pm->mesh;
    pm->kblocks = grid_for(m.gc, kBlock, 4096);
    if (st == HYMD_OK) st = pm->paint.reserve(sizeof(int64_t) * m.g * max_types);
    if (st == HYMD_OK) st = pm->spec.reserve(sizeof(float2) * m.gc * max_types);
    if (st == HYMD_OK) st = pm->grad.reserve(sizeof(float2) * m.gc * std::max(3 * max_types, 4));
    // energy partial rows: k-space blocks or x-line blocks (at least 2 lines each)
    if (st == HYMD_OK)
        st = pm->partials.reserve(sizeof(double) * (std::max<int64_t>(pm->kblocks, (int64_t)m.my * m.zh / 2 + 1) + 16));
    if (st == HYMD_OK) st = pm->result.reserve(sizeof(double) * 16);
    if (st == HYMD_OK) st = pm->host.reserve(sizeof(double) * 16);
    if (st == HYMD_OK) st = pm->counts.reserve(sizeof(int) * (m.g + 2));
    if (st == HYMD_OK) st = pm->start.reserve(sizeof(int) * (m.g + 1));
    if (st == HYMD_OK) st = pm->tiles.reserve(sizeof(int) * (scan_tiles(m.g) + 1));
    if (st == HYMD_OK) {
        zero_kernel<<<grid_for(m.g * max_types), kBlock, 0, pm->stream>>>(
            (unsigned long long *)pm->paint.ptr, m.g * max_types);
        zero_kernel<<<grid_for(m.g / 2 + 1), kBlock, 0, pm->stream>>>((unsigned long long *)pm->counts.ptr,
                                                                       m.g / 2 + 1);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(pm->stream);
        if (e != hipSuccess) {
            set_error("zeroing the paint grid failed: %s", hipGetErrorString(e));
            st = HYMD_ERR_HIP;
        }
    }
    if (st != HYMD_OK) {
        delete pm;
        return st;
    }
    *out = pm;
    return HYMD_OK;
}

extern "C" int hymd_pm_destroy(hymd_pm *pm) {
    if (!pm) return HYMD_OK;
    (void)hipStreamSynchronize(pm->stream);
    delete pm;
    return HYMD_OK;
}

extern "C" int hymd_pm_set_stream(hymd_pm *pm, void *stream) {
    HYMD_REQUIRE(pm, "null handle");
    pm->stream = (hipStream_t)stream;
    return HYMD_OK;
}

extern "C" int hymd_pm_set_tuning(hymd_pm *pm, int32_t knob, int32_t value) {
    HYMD_REQUIRE(pm, "null handle");
    switch (knob) {
    case HYMD_TUNE_PAINT_ROWS:
        HYMD_REQUIRE(value == 1 || value == kPaintWaveRows, "paint rows per workgroup must be 1 or %d", kPaintWaveRows);
        pm->paint_rows = value;
        break;
    case HYMD_TUNE_READOUT_BLOCK:
        HYMD_REQUIRE(value == 64 || value == 128 || value == 256, "readout block must be 64, 128 or 256");
        pm->readout_block = value;
        break;
    case HYMD_TUNE_XLINE_WIDTH:
        HYMD_REQUIRE(value == 0 || value == 2 || value == 4 || value == 8 || value == 16,
                     "x-line width must be 0, 2, 4, 8 or 16");
        pm->xline_w = value;
        break;
    default:
        set_error("unknown tuning knob %d", knob);
        return HYMD_ERR_INVALID;
    }
    ++pm->box_version;  // captured steps replay the old launch shapes: capture again
    return HYMD_OK;
}

extern "C" int hymd_pm_get_tuning(hymd_pm *pm, int32_t knob, int32_t *value) {
    HYMD_REQUIRE(pm && value, "null argument");
    switch (knob) {
    case HYMD_TUNE_PAINT_ROWS: *value = pm->paint_rows; break;
    case HYMD_TUNE_READOUT_BLOCK: *value = pm->readout_block; break;
    case HYMD_TUNE_XLINE_WIDTH: *value = pm->xline_w; break;
    default:
        set_error("unknown tuning knob %d", knob);
        return HYMD_ERR_INVALID;
    }
    return HYMD_OK;
}

extern "C" int hymd_pm_set_box(hymd_pm *pm, const double box[3]) {
    HYMD_REQUIRE(pm && box, "null argument");
    int32_t mesh[3] = {pm->mesh.mx, pm->mesh.my, pm->mesh.mz};
    HYMD_TRY(set_mesh(pm, mesh, box));
    ++pm->box_version;
    return HYMD_OK;
}

static FieldK field_params(int T, const double *coupling, const double *shift, double constant, double sigma,
                           const Mesh &m) {
    FieldK p;
    p.T = T;
    p.want_energy = 0;
    p.sigma2 = sigma * sigma;
    p.inv_g = 1.0 / (double)m.g;
    p.c0 = constant;
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) {
        p.b[t] = (shift && t < T) ? shift[t] : 0.0;
        for (int u = 0; u < HYMD_MAX_TYPES; ++u) p.Ad[t * HYMD_MAX_TYPES + u] = (t < T && u < T) ? coupling[t * T + u] : 0.0;
    }
    return p;
}

// Bin particles by the mesh row of their CIC base cell: perm[pos] = original
// index, xs/ts/ws the sorted positions, types (0 when types == NULL) and
// weights (if w != NULL), start[r] .. start[r+1] the particles based in row r.
// Bins the n particles by mesh row.  keys_ready: the keys (and counts) were
// produced by field_key_target()'s producer.
static int sort_particles(hymd_pm *pm, const float *x, const int32_t *types, const float *w, int64_t n,
                          bool keys_ready = false) {
    const Mesh &m = pm->mesh;
    hipStream_t s = pm->stream;
    HYMD_REQUIRE(n < (int64_t(1) << 31), "too many particles to sort (%lld)", (long long)n);
    const int64_t nrows = (int64_t)m.mx * m.my;
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    HYMD_TRY(pm->keys.reserve(sizeof(int2) * nn));
    HYMD_TRY(pm->perm.reserve(sizeof(int) * nn));
    HYMD_TRY(pm->xs.reserve(sizeof(float4) * nn));
    if (w) HYMD_TRY(pm->ws.reserve(sizeof(float) * nn));
    if (n > 0 && !keys_ready)
        row_key_kernel<<<grid_for(n), kBlock, 0, s>>>(x, w, n, m, (int2 *)pm->keys.ptr, (int *)pm->counts.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan((int *)pm->counts.ptr, (int *)pm->start.ptr, nrows, (int *)pm->tiles.ptr, nullptr,
                            true, s));
    HYMD_TRY(launch_key_scatter(x, types, w, n, (const int2 *)pm->keys.ptr, (const int *)pm->start.ptr,
                                (int *)pm->perm.ptr, (float4 *)pm->xs.ptr, w ? (float *)pm->ws.ptr : nullptr, s));
    pm->sorted_n = n;
    pm->sorted_weights = (w != nullptr);
    pm->density_spectra = false;
    return HYMD_OK;
}

// Row paint of the binned particles (weights ws per binned slot, or the
// per-type masses) into T float grids sum * out_scale, optionally also the
// int64 sums; one workgroup or one wavefront per mesh row (pm->paint_rows).
static int launch_row_paint(hymd_pm *pm, int T, const float *ws, const PaintArgs &pa, double out_scale, float *out,
                            long long *out_fixed) {
    const Mesh &m = pm->mesh;
    const int64_t nrows = (int64_t)m.mx * m.my;
    if (pm->paint_rows == kPaintWaveRows) {
        const int tc = std::max(1, std::min(T, (int)(kRowLdsBytes * 2 / (sizeof(long long) * kPaintWaveRows * m.mz))));
        dim3 grid((unsigned)((nrows + kPaintWaveRows - 1) / kPaintWaveRows), (unsigned)((T + tc - 1) / tc));
        paint_rows_wave_kernel<<<grid, kBlock, sizeof(long long) * kPaintWaveRows * tc * m.mz, pm->stream>>>(
            (const float4 *)pm->xs.ptr, ws, (const int *)pm->start.ptr, m, pa, tc, out_scale, out, out_fixed);
    } else {
        const int tc = std::max(1, std::min(T, (int)(kRowLdsBytes / (sizeof(long long) * m.mz))));
        dim3 grid((unsigned)nrows, (unsigned)((T + tc - 1) / tc));
        paint_rows_kernel<<<grid, kBlock, sizeof(long long) * tc * m.mz, pm->stream>>>(
            (const float4 *)pm->xs.ptr, ws, (const int *)pm->start.ptr, m, pa, tc, out_scale, out, out_fixed);
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// Row-gather paint of the sorted particles into float grids rho_t = paint/dV
// (T grids), optionally also the int64 fixed-point sums.
static int paint_sorted(hymd_pm *pm, int T, const double *mass, float *out, long long *out_fixed) {
    PaintArgs pa;
    pa.n_types = T;
    pa.scale = std::ldexp(1.0, HYMD_FIXED_SHIFT);
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) pa.mass[t] = (mass && t < T) ? mass[t] : 1.0;
    const double out_scale = std::ldexp(1.0, -HYMD_FIXED_SHIFT) / cell_volume(pm);
    return launch_row_paint(pm, T, pm->sorted_weights ? (const float *)pm->ws.ptr : nullptr, pa, out_scale, out,
                            out_fixed);
}

// Readout of ncomp grids per type at the last sorted particle set, written to
// the particles' original rows (scaled by their weights if the sort had them);
// accumulate: out += the readout (rows the binning left out keep their
// values) instead of out = it.
static int readout_sorted(hymd_pm *pm, int ncomp, int n_types, const float *grids, float *out, bool weighted,
                          bool accumulate = false) {
    const int64_t n = pm->sorted_n;
    if (n > 0) {
        const float4 *rec = (const float4 *)pm->xs.ptr;
        const float *ws = weighted ? (const float *)pm->ws.ptr : nullptr;
        const int *pp = (const int *)pm->perm.ptr;
        // a weighted binning leaves zero-weight particles out: their rows get
        // zero, the kept ones are counted in start[rows]
        const int *kept = nullptr;
        if (pm->sorted_weights) {
            if (!accumulate) HYMD_TRY(zero_words(out, ncomp * n, pm->stream));
            kept = (const int *)pm->start.ptr + (int64_t)pm->mesh.mx * pm->mesh.my;
        }
        const int bs = pm->readout_block;
        // a weighted binning keeps a count known on the device only: a grid
        // of at most 4096 blocks loops over it
        const int nb = kept ? grid_for(n, bs, 4096) : grid_for(n, bs);
        const int acc = accumulate ? 1 : 0;
        if (kept) {
            if (ncomp == 3)
                readout_kernel<3, false><<<nb, bs, 0, pm->stream>>>(grids, rec, nullptr, ws, n, n_types, pm->mesh,
                                                                    out, pp, kept, acc, 0, 0);
            else
                readout_kernel<1, false><<<nb, bs, 0, pm->stream>>>(grids, rec, nullptr, ws, n, n_types, pm->mesh,
                                                                    out, pp, kept, acc, 0, 0);
        } else {
            // z lines of 32 floats (one 128-byte cache line), coarser when
            // the (type, line) buckets would exceed kSortBuckets
            int zshift = 0;
            while (n_types * (((pm->mesh.mz - 1) >> (5 + zshift)) + 1) > kSortBuckets) ++zshift;
            const int nzl = ((pm->mesh.mz - 1) >> (5 + zshift)) + 1;
            if (ncomp == 3)
                readout_kernel<3, true><<<nb, bs, 0, pm->stream>>>(grids, rec, nullptr, ws, n, n_types, pm->mesh, out,
                                                                   pp, nullptr, acc, zshift, nzl);
            else
                readout_kernel<1, true><<<nb, bs, 0, pm->stream>>>(grids, rec, nullptr, ws, n, n_types, pm->mesh, out,
                                                                   pp, nullptr, acc, zshift, nzl);
        }
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

namespace hymd {
// compute_field_force at the particles painted by the last update_field on
// this handle (positions and types unchanged since)
int field_readout_last(hymd_pm *pm, int n_types, const float *force_mesh, float *force) {
    return readout_sorted(pm, 3, n_types, force_mesh, force, false);
}
}  // namespace hymd

static int launch_paint(hymd_pm *pm, const float *x, const int32_t *types, const float *weights, int64_t n,
                        int32_t n_types, const double *mass, unsigned long long *grid) {
    PaintArgs pa;
    pa.n_types = n_types;
    pa.scale = std::ldexp(1.0, HYMD_FIXED_SHIFT);
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) pa.mass[t] = (mass && t < n_types) ? mass[t] : 1.0;
    if (n > 0)
        paint_fixed_kernel<float><<<grid_for(n), kBlock, 0, pm->stream>>>(x, types, weights, n, pm->mesh, pa,
                                                                            grid);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

static int read_result(hymd_pm *pm, int nblocks, int nv, double *out) {
    HYMD_TRY(reduce_partials((const double *)pm->partials.ptr, nblocks, nv, (double *)pm->result.ptr, pm->stream));
    HYMD_HIP(hipMemcpyAsync(pm->host.ptr, pm->result.ptr, sizeof(double) * nv, hipMemcpyDeviceToHost, pm->stream));
    HYMD_HIP(hipStreamSynchronize(pm->stream));
    for (int j = 0; j < nv; ++j) out[j] = ((double *)pm->host.ptr)[j];
    return HYMD_OK;
}

extern "C" int hymd_paint_fixed(hymd_pm *pm, const float *x, const int32_t *types, const float *weights,
                                int64_t n, int32_t n_types, const double *mass, int64_t *grid,
                                int32_t accumulate) {
    HYMD_REQUIRE(pm && grid && (x || n == 0), "null argument");
    HYMD_REQUIRE(n >= 0, "negative particle count");
    HYMD_REQUIRE(n_types >= 1 && n_types <= pm->max_types, "n_types out of range");
    const int ngrids = types ? n_types : 1;
    if (!accumulate) {
        // the hot path: cell sort + gather (int64 sums written to grid)
        HYMD_TRY(sort_particles(pm, x, types, weights, n));
        return paint_sorted(pm, ngrids, mass, (float *)pm->paint.ptr, (long long *)grid);
    }
    // accumulate into the caller's grid with one 64-bit atomic per corner
    return launch_paint(pm, x, types, weights, n, ngrids, mass, (unsigned long long *)grid);
}

extern "C" int hymd_fft_r2c(hymd_pm *pm, const float *in, void *out, int32_t batch) {
    HYMD_REQUIRE(pm && in && out && batch >= 1, "bad argument");
    return fft_r2c_real(pm->fft, in, (float2 *)out, batch, pm->stream);
}

extern "C" int hymd_fft_c2r(hymd_pm *pm, const void *in, float *out, int32_t batch) {
    HYMD_REQUIRE(pm && in && out && batch >= 1, "bad argument");
    const size_t bytes = sizeof(float2) * pm->mesh.gc * batch;
    HYMD_TRY(pm->cscratch.reserve(bytes));
    HYMD_HIP(hipMemcpyAsync(pm->cscratch.ptr, in, bytes, hipMemcpyDeviceToDevice, pm->stream));
    return fft_c2r(pm->fft, (float2 *)pm->cscratch.ptr, out, batch, 1.f, pm->stream);
}

extern "C" int hymd_readout(hymd_pm *pm, const float *grid, const float *x, int64_t n, float *out) {
    HYMD_REQUIRE(pm && grid && out && (x || n == 0) && n >= 0, "bad argument");
    if (n > 0)
        readout_kernel<1, false><<<grid_for(n), kBlock, 0, pm->stream>>>(grid, nullptr, x, nullptr, n, 1, pm->mesh,
                                                                           out, nullptr, nullptr, 0, 0, 0);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

static int update_field_impl(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                             const double *coupling, const double *shift, double constant, const double *mass,
                             double sigma, float *force_mesh, float *phi, void *phi_fourier, float *v_ext,
                             double *energy, bool keys_ready) {
    HYMD_REQUIRE(pm && force_mesh && coupling && (x || n == 0) && (types || n == 0), "null argument");
    HYMD_REQUIRE(n >= 0, "negative particle count");
    HYMD_REQUIRE(n_types >= 1 && n_types <= pm->max_types, "n_types %d exceeds the handle's %d", n_types,
                 pm->max_types);
    HYMD_REQUIRE(sigma >= 0.0 && std::isfinite(sigma), "sigma must be finite and >= 0");
    const Mesh &m = pm->mesh;
    const int T = n_types;
    hipStream_t s = pm->stream;
    unsigned long long *grid = (unsigned long long *)pm->paint.ptr;
    float2 *spec = (float2 *)pm->spec.ptr;
    float2 *grad = (float2 *)pm->grad.ptr;

    // 1. cell sort + gather paint of rho_t = paint_t / dV (float grids)
    float *rho = (float *)grid;  // the handle's paint buffer holds T float grids here
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n, keys_ready));
    HYMD_TRY(paint_sorted(pm, T, mass, rho, nullptr));
    FieldK p = field_params(T, coupling, shift, constant, sigma, m);
    p.want_energy = energy ? 1 : 0;
    KGeom g0 = kgeom(pm);
    const XlineShape sh = T <= 8 ? xline_grad_shape(pm->fft, T, pm->xline_w) : XlineShape{};
    if (!phi && !phi_fourier && !v_ext && sh.W > 0) {
        // force grids only: z + y forward passes; one pass for the forward x
        // transform, the k-space pass and the 2T inverse x transforms; the
        // y-z plane pass writes the 3T force grids
        HYMD_TRY(fft_r2c_zy(pm->fft, rho, spec, T, s));
        double *part = energy ? (double *)pm->partials.ptr : nullptr;
        int nb = 0;
        HYMD_TRY(launch_xline_grad(pm, spec, grad, g0, p, sh, part, &nb));
        if (pm->fft.lds_plane) {
            HYMD_TRY(fft_c2r_yz_grad(pm->fft, grad, force_mesh, T, g0.kfac[1], g0.kfac[2], 1.f, s));
        } else {  // y-z planes too large for LDS: the y and z line passes
            HYMD_TRY(pm->cscratch.reserve(sizeof(float2) * m.gc * 3 * T));
            HYMD_TRY(fft_c2r_lines_grad(pm->fft, grad, (float2 *)pm->cscratch.ptr, force_mesh, T, g0.kfac[1],
                                        g0.kfac[2], 1.f, s));
        }
        if (energy) {
            double e = 0.0;
            HYMD_TRY(read_result(pm, nb, 1, &e));
            *energy = e * (m.box[0] * m.box[1] * m.box[2]);
        }
        return HYMD_OK;
    }
    const int W = xline_width(pm->fft, T);
    if (!phi && !phi_fourier && !v_ext && W > 0) {
        // force grids only: z + y forward passes, the forward x pass fused
        // with the k-space pass, then the inverse x, y and z passes
        HYMD_TRY(fft_r2c_zy(pm->fft, rho, spec, T, s));
        double *part = energy ? (double *)pm->partials.ptr : nullptr;
        int nb = 0;
        if (T <= 4) HYMD_TRY(launch_xline<4>(pm, spec, grad, g0, p, W, part, &nb));
        else if (T <= 8) HYMD_TRY(launch_xline<8>(pm, spec, grad, g0, p, W, part, &nb));
        else HYMD_TRY(launch_xline<16>(pm, spec, grad, g0, p, W, part, &nb));
        HYMD_TRY(fft_c2r(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
        if (energy) {
            double e = 0.0;
            HYMD_TRY(read_result(pm, nb, 1, &e));
            *energy = e * (m.box[0] * m.box[1] * m.box[2]);
        }
        return HYMD_OK;
    }
    // 2. r2c
    HYMD_TRY(fft_r2c_real(pm->fft, rho, spec, T, s));
    // 3. fused filter / W' / filter / gradient
    float2 *phif = (float2 *)phi_fourier;
    if (phi && !phif) {
        HYMD_TRY(pm->cscratch.reserve(sizeof(float2) * m.gc * T));
        phif = (float2 *)pm->cscratch.ptr;
    }
    KGeom g = kgeom(pm);
    double *partials = energy ? (double *)pm->partials.ptr : nullptr;
    const int kb = pm->kblocks;
    if (T <= 4)
        field_kspace_kernel<4><<<kb, kBlock, 0, s>>>(spec, grad, phif, v_ext ? 1 : 0, g, p, partials);
    else if (T <= 8)
        field_kspace_kernel<8><<<kb, kBlock, 0, s>>>(spec, grad, phif, v_ext ? 1 : 0, g, p, partials);
    else
        field_kspace_kernel<16><<<kb, kBlock, 0, s>>>(spec, grad, phif, v_ext ? 1 : 0, g, p, partials);
    HYMD_HIP(hipGetLastError());
    // 4. force grids
    HYMD_TRY(fft_c2r(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
    // 5. optional real-space outputs
    if (v_ext) {
        HYMD_TRY(fft_c2r(pm->fft, spec, v_ext, T, 1.f, s));
    }
    if (phi) {
        float2 *work = phif;
        if (phi_fourier) {  // keep the caller's spectra intact
            HYMD_TRY(pm->cscratch.reserve(sizeof(float2) * m.gc * T));
            HYMD_HIP(hipMemcpyAsync(pm->cscratch.ptr, phi_fourier, sizeof(float2) * m.gc * T,
                                    hipMemcpyDeviceToDevice, s));
            work = (float2 *)pm->cscratch.ptr;
        }
        HYMD_TRY(fft_c2r(pm->fft, work, phi, T, 1.f, s));
    }
    if (energy) {
        double e = 0.0;
        HYMD_TRY(read_result(pm, kb, 1, &e));
        *energy = e * (m.box[0] * m.box[1] * m.box[2]);
    }
    return HYMD_OK;
}

extern "C" int hymd_update_field(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                                 const double *coupling, const double *shift, double constant,
                                 const double *mass, double sigma, float *force_mesh, float *phi,
                                 void *phi_fourier, float *v_ext, double *energy) {
    return update_field_impl(pm, x, types, n, n_types, coupling, shift, constant, mass, sigma, force_mesh, phi,
                             phi_fourier, v_ext, energy, false);
}

namespace hymd {
int field_key_target(hymd_pm *pm, int64_t n, RowKeyTarget *t) {
    HYMD_REQUIRE(pm && t && n >= 0 && n < (int64_t(1) << 31), "bad argument");
    HYMD_TRY(pm->keys.reserve(sizeof(int2) * (size_t)std::max<int64_t>(n, 1)));
    t->keys = (int2 *)pm->keys.ptr;
    t->counts = (int *)pm->counts.ptr;
    t->inv_h[0] = pm->mesh.inv_h[0];
    t->inv_h[1] = pm->mesh.inv_h[1];
    t->mx = pm->mesh.mx;
    t->my = pm->mesh.my;
    return HYMD_OK;
}

int update_field_keyed(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                       const double *coupling, const double *shift, double constant, const double *mass,
                       double sigma, float *force_mesh, double *energy, void *phi_fourier) {
    return update_field_impl(pm, x, types, n, n_types, coupling, shift, constant, mass, sigma, force_mesh, nullptr,
                             phi_fourier, nullptr, energy, true);
}
}  // namespace hymd

__global__ void __launch_bounds__(kBlock) row_order_kernel(const int2 *__restrict__ keys, int64_t n,
                                                           const int *__restrict__ start, int *__restrict__ order) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int2 kr = keys[i];
        order[start[kr.x] + kr.y] = (int)i;
    }
}

namespace hymd {
// order[slot] = i: the n particles at x (float32 rows) in mesh-row order (the
// binning of update_field, without its records).  Uses the handle's binning
// buffers, so it runs between field updates.
int row_order(hymd_pm *pm, const float *x, int64_t n, int *order, hipStream_t s) {
    HYMD_REQUIRE(pm && (x || n == 0) && order && n >= 0 && n < (int64_t(1) << 31), "bad argument");
    if (n == 0) return HYMD_OK;
    const Mesh &m = pm->mesh;
    HYMD_TRY(pm->keys.reserve(sizeof(int2) * (size_t)n));
    row_key_kernel<<<grid_for(n), kBlock, 0, s>>>(x, nullptr, n, m, (int2 *)pm->keys.ptr, (int *)pm->counts.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan((int *)pm->counts.ptr, (int *)pm->start.ptr, (int64_t)m.mx * m.my, (int *)pm->tiles.ptr,
                            nullptr, true, s));
    row_order_kernel<<<grid_for(n), kBlock, 0, s>>>((const int2 *)pm->keys.ptr, n, (const int *)pm->start.ptr, order);
    HYMD_HIP(hipGetLastError());
    pm->sorted_n = 0;  // the last binned set is gone
    pm->density_spectra = false;
    return HYMD_OK;
}
}  // namespace hymd

// paint + r2c + moments (and, if force_mesh != NULL, the force grids of the
// quadratic Hamiltonian (coupling, shift) as in hymd_update_field)
extern "C" int hymd_field_moments(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                                  const double *mass, double sigma, double *moments, const double *coupling,
                                  const double *shift, float *force_mesh) {
    HYMD_REQUIRE(pm && moments && (x || n == 0) && (types || n == 0) && n >= 0, "bad argument");
    HYMD_REQUIRE(n_types >= 1 && n_types <= 8 && n_types <= pm->max_types, "moments support n_types <= 8");
    HYMD_REQUIRE(!force_mesh || coupling, "force grids need the coupling matrix");
    const Mesh &m = pm->mesh;
    const int T = n_types;
    hipStream_t s = pm->stream;
    float *rho = (float *)pm->paint.ptr;
    float2 *spec = (float2 *)pm->spec.ptr;
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
    HYMD_TRY(paint_sorted(pm, T, mass, rho, nullptr));
    HYMD_TRY(fft_r2c_real(pm->fft, rho, spec, T, s));
    KGeom g = kgeom(pm);
    const int np = T * (T + 1) / 2;
    const int nv = 2 * np + T;
    const int row = moments_row_len(T);
    const int nb = moments_blocks(m.gc);
    DevBuf &part = pm->mpart;
    HYMD_TRY(part.reserve(sizeof(double) * (size_t)row * nb));
    HYMD_TRY(pm->result.reserve(sizeof(double) * row));
    HYMD_TRY(pm->host.reserve(sizeof(double) * row));
    if (T <= 2)
        field_moments_kernel<2><<<nb, kBlock, 0, s>>>(spec, g, T, sigma * sigma, 1.0 / (double)m.g,
                                                      (double *)part.ptr);
    else if (T <= 4)
        field_moments_kernel<4><<<nb, kBlock, 0, s>>>(spec, g, T, sigma * sigma, 1.0 / (double)m.g,
                                                      (double *)part.ptr);
    else
        field_moments_kernel<8><<<nb, kBlock, 0, s>>>(spec, g, T, sigma * sigma, 1.0 / (double)m.g,
                                                      (double *)part.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials((const double *)part.ptr, nb, row, (double *)pm->result.ptr, s));
    HYMD_HIP(hipMemcpyAsync(pm->host.ptr, pm->result.ptr, sizeof(double) * row, hipMemcpyDeviceToHost, s));
    if (force_mesh) {
        FieldK p = field_params(T, coupling, shift, 0.0, sigma, m);
        float2 *grad = (float2 *)pm->grad.ptr;
        if (T <= 4)
            field_kspace_kernel<4><<<pm->kblocks, kBlock, 0, s>>>(spec, grad, nullptr, 0, g, p, nullptr);
        else
            field_kspace_kernel<8><<<pm->kblocks, kBlock, 0, s>>>(spec, grad, nullptr, 0, g, p, nullptr);
        HYMD_HIP(hipGetLastError());
        HYMD_TRY(fft_c2r(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
    }
    HYMD_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nv; ++j) moments[j] = ((double *)pm->host.ptr)[j];
    pm->density_spectra = true;
    pm->density_types = T;
    return HYMD_OK;
}

// Force grids (and the forces of the binned particles) of a quadratic
// Hamiltonian from the density spectra held by the last hymd_field_moments
// call: the k-space pass and the inverse transforms, no binning or paint.
extern "C" int hymd_field_forces_held(hymd_pm *pm, int32_t n_types, const double *coupling, const double *shift,
                                      double sigma, float *force_mesh, float *force) {
    HYMD_REQUIRE(pm && coupling && force_mesh, "null argument");
    HYMD_REQUIRE(sigma >= 0.0 && std::isfinite(sigma), "sigma must be finite and >= 0");
    HYMD_REQUIRE(pm->density_spectra && pm->density_types == n_types,
                 "no density spectra of %d types are held (call hymd_field_moments first)", n_types);
    const Mesh &m = pm->mesh;
    const int T = n_types;
    hipStream_t s = pm->stream;
    FieldK p = field_params(T, coupling, shift, 0.0, sigma, m);
    float2 *grad = (float2 *)pm->grad.ptr;
    KGeom g = kgeom(pm);
    if (T <= 4)
        field_kspace_kernel<4><<<pm->kblocks, kBlock, 0, s>>>((float2 *)pm->spec.ptr, grad, nullptr, 0, g, p, nullptr);
    else
        field_kspace_kernel<8><<<pm->kblocks, kBlock, 0, s>>>((float2 *)pm->spec.ptr, grad, nullptr, 0, g, p, nullptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(fft_c2r(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
    if (force) HYMD_TRY(readout_sorted(pm, 3, T, force_mesh, force, false));
    return HYMD_OK;
}

// Adjoint of hymd_compute_field_force o hymd_update_field (coupling A,
// sigma): dL/dA and dL/dsigma for the upstream gradient grad_force = dL/dF.
extern "C" int hymd_field_adjoint(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                                  const double *mass, double sigma, const double *coupling,
                                  const float *grad_force, int32_t reuse, double *d_coupling, double *d_sigma) {
    HYMD_REQUIRE(pm && coupling && d_coupling && (grad_force || n == 0) && n >= 0, "bad argument");
    HYMD_REQUIRE(n_types >= 1 && n_types <= 8 && n_types <= pm->max_types, "the adjoint supports n_types <= 8");
    HYMD_REQUIRE(sigma >= 0.0 && std::isfinite(sigma), "sigma must be finite and >= 0");
    const Mesh &m = pm->mesh;
    const int T = n_types;
    hipStream_t s = pm->stream;
    float2 *spec = (float2 *)pm->spec.ptr;
    if (reuse) {
        HYMD_REQUIRE(pm->density_spectra && pm->density_types == T && pm->sorted_n == n && !pm->sorted_weights,
                     "reuse needs the binning and density spectra of a hymd_field_moments call on the same "
                     "particles and types");
    } else {
        HYMD_REQUIRE((x || n == 0) && (types || n == 0), "null positions or types");
        HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
        HYMD_TRY(paint_sorted(pm, T, mass, (float *)pm->paint.ptr, nullptr));
        HYMD_TRY(fft_r2c_real(pm->fft, (const float *)pm->paint.ptr, spec, T, s));
        pm->density_spectra = true;
        pm->density_types = T;
    }
    for (int j = 0; j < T * T; ++j) d_coupling[j] = 0.0;
    if (d_sigma) *d_sigma = 0.0;
    if (n == 0) return HYMD_OK;
    // 1. upstream gradient into binned order, and its largest modulus
    HYMD_TRY(pm->adj.reserve(sizeof(float) * (3 * (size_t)n + 64)));
    float *ws = (float *)pm->adj.ptr;
    unsigned *maxbits = (unsigned *)(ws + 3 * n);
    HYMD_TRY(zero_words(maxbits, 1, s));
    adjoint_gather_kernel<<<grid_for(n, kBlock, 4096), kBlock, 0, s>>>(grad_force, (const int *)pm->perm.ptr, n, ws,
                                                                       maxbits);
    HYMD_HIP(hipGetLastError());
    unsigned hbits = 0;
    HYMD_HIP(hipMemcpyAsync(&hbits, maxbits, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    float gmax;
    std::memcpy(&gmax, &hbits, sizeof(float));
    HYMD_REQUIRE(std::isfinite(gmax), "grad_force holds a non-finite value");
    if (gmax == 0.f) return HYMD_OK;
    // 2. CIC paint of each component per type (the readout's adjoint), as
    // int64 fixed point scaled so that the largest deposit is ~2^36
    int ex = 0;
    (void)std::frexp((double)gmax, &ex);
    PaintArgs pa;
    pa.n_types = T;
    pa.scale = std::ldexp(1.0, 36 - ex);
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) pa.mass[t] = 1.0;
    HYMD_TRY(pm->rscratch.reserve(sizeof(float) * 3 * (size_t)T * m.g));
    float *pg = (float *)pm->rscratch.ptr;
    for (int d = 0; d < 3; ++d)
        HYMD_TRY(launch_row_paint(pm, T, ws + d * n, pa, 1.0 / pa.scale, pg + (int64_t)d * T * m.g, nullptr));
    // 3. their spectra, 4. the adjoint moments
    float2 *pspec = (float2 *)pm->grad.ptr;
    HYMD_TRY(fft_r2c_real(pm->fft, pg, pspec, 3 * T, s));
    KGeom g = kgeom(pm);
    AdjK ak{};
    for (int t = 0; t < T; ++t)
        for (int u = 0; u < T; ++u) ak.A[t][u] = coupling[t * T + u];
    const int nb = moments_blocks(m.gc);
    const int mt = T <= 2 ? 2 : (T <= 4 ? 4 : 8);
    const int nv = mt * mt + 1;
    HYMD_TRY(pm->mpart.reserve(sizeof(double) * (size_t)nv * nb));
    HYMD_TRY(pm->result.reserve(sizeof(double) * nv));
    HYMD_TRY(pm->host.reserve(sizeof(double) * nv));
    double *part = (double *)pm->mpart.ptr;
    if (mt == 2)
        adjoint_moments_kernel<2><<<nb, kBlock, 0, s>>>(spec, pspec, g, T, sigma * sigma, ak, part);
    else if (mt == 4)
        adjoint_moments_kernel<4><<<nb, kBlock, 0, s>>>(spec, pspec, g, T, sigma * sigma, ak, part);
    else
        adjoint_moments_kernel<8><<<nb, kBlock, 0, s>>>(spec, pspec, g, T, sigma * sigma, ak, part);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials(part, nb, nv, (double *)pm->result.ptr, s));
    HYMD_HIP(hipMemcpyAsync(pm->host.ptr, pm->result.ptr, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    const double *r = (const double *)pm->host.ptr;
    const double inv_g = 1.0 / (double)m.g;
    for (int t = 0; t < T; ++t)
        for (int u = 0; u < T; ++u) d_coupling[t * T + u] = r[t * mt + u] * inv_g;
    if (d_sigma) *d_sigma = -2.0 * sigma * r[mt * mt] * inv_g;
    return HYMD_OK;
}

extern "C" int hymd_field_pressure(hymd_pm *pm, const void *phi_fourier, int32_t n_types, const double *coupling,
                                   const double *shift, double constant, double sigma, const float *psi,
                                   const float *phi_q, const double *type_charges, double self_energy,
                                   double *out) {
    HYMD_REQUIRE(pm && phi_fourier && coupling && shift && out, "null argument");
    HYMD_REQUIRE(n_types >= 1 && n_types <= 8, "the pressure supports n_types <= 8");
    HYMD_REQUIRE(!psi || type_charges, "psi needs the per-type charges");
    HYMD_REQUIRE(!phi_q || psi, "phi_q enters with psi only");
    const Mesh &m = pm->mesh;
    const int T = n_types;
    hipStream_t s = pm->stream;
    PressK p{};
    for (int t = 0; t < T; ++t) {
        for (int u = 0; u < T; ++u) p.A[t][u] = coupling[t * T + u];
        p.q[t] = psi ? type_charges[t] : 0.0;
    }
    p.inv_g = 1.0 / (double)m.g;
    float2 *psif = nullptr, *qf = nullptr;
    if (psi) {  // raw spectra of psi and phi_q into the gradient scratch (>= 4 spectra)
        psif = (float2 *)pm->grad.ptr;
        HYMD_TRY(fft_r2c_real(pm->fft, psi, psif, 1, s));
        if (phi_q) {
            qf = psif + m.gc;
            HYMD_TRY(fft_r2c_real(pm->fft, phi_q, qf, 1, s));
        }
    }
    KGeom g = kgeom(pm);
    const int nb = moments_blocks(m.gc);
    DevBuf &part = pm->mpart;
    HYMD_TRY(part.reserve(sizeof(double) * (size_t)kPressNv * nb));
    HYMD_TRY(pm->result.reserve(sizeof(double) * kPressNv));
    HYMD_TRY(pm->host.reserve(sizeof(double) * kPressNv));
    const float2 *phif = (const float2 *)phi_fourier;
    if (T <= 2)
        field_pressure_kernel<2><<<nb, kBlock, 0, s>>>(phif, psif, qf, g, T, p, (double *)part.ptr);
    else if (T <= 4)
        field_pressure_kernel<4><<<nb, kBlock, 0, s>>>(phif, psif, qf, g, T, p, (double *)part.ptr);
    else
        field_pressure_kernel<8><<<nb, kBlock, 0, s>>>(phif, psif, qf, g, T, p, (double *)part.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials((const double *)part.ptr, nb, kPressNv, (double *)pm->result.ptr, s));
    HYMD_HIP(hipMemcpyAsync(pm->host.ptr, pm->result.ptr, sizeof(double) * kPressNv, hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    const double *r = (const double *)pm->host.ptr;
    double bphi = 0.0;
    for (int t = 0; t < T; ++t) bphi += shift[t] * r[9 + t];
    const double V = m.box[0] * m.box[1] * m.box[2];
    // p0 = -(dV/V) sum_cells (w + w_elec), p1 = (dV/V) sum V_bar phi~, p2_d = sigma^2 (dV/V) sum V_bar lap_d phi~
    double p0 = -(0.5 * r[0] + bphi + constant);
    if (phi_q) p0 += -0.5 * r[8] + self_energy / V;
    const double p1 = r[0] + bphi + r[4];
    out[0] = p0;
    out[1] = p1;
    for (int d = 0; d < 3; ++d) out[2 + d] = -sigma * sigma * (r[1 + d] + r[5 + d]);
    return HYMD_OK;
}

extern "C" int hymd_compute_field_force(hymd_pm *pm, const float *x, const int32_t *types, int64_t n,
                                        int32_t n_types, const float *force_mesh, float *force) {
    HYMD_REQUIRE(pm && force_mesh && force && (x || n == 0) && (types || n == 0), "null argument");
    HYMD_REQUIRE(n >= 0 && n_types >= 1 && n_types <= HYMD_MAX_TYPES, "bad sizes");
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
    return readout_sorted(pm, 3, n_types, force_mesh, force, false);
}

namespace hymd {
int update_field_force_q_impl(hymd_pm *pm, const float *x, const float *charges, int64_t n, double sigma,
                              double elec_factor, double self_energy, float *elec_forces, float *phi_q, float *psi,
                              float *elec_field, double *energy_q, bool accumulate);
}

extern "C" int hymd_update_field_force_q(hymd_pm *pm, const float *x, const float *charges, int64_t n,
                                         double sigma, double elec_factor, double self_energy,
                                         float *elec_forces, float *phi_q, float *psi, float *elec_field,
                                         double *energy_q) {
    return update_field_force_q_impl(pm, x, charges, n, sigma, elec_factor, self_energy, elec_forces, phi_q, psi,
                                     elec_field, energy_q, false);
}

int hymd::update_field_force_q_impl(hymd_pm *pm, const float *x, const float *charges, int64_t n, double sigma,
                                    double elec_factor, double self_energy, float *elec_forces, float *phi_q,
                                    float *psi, float *elec_field, double *energy_q, bool accumulate) {
    HYMD_REQUIRE(pm && elec_forces && (n == 0 || (x && charges)) && n >= 0, "bad argument");
    const Mesh &m = pm->mesh;
    hipStream_t s = pm->stream;
    float2 *spec = (float2 *)pm->spec.ptr;
    float2 *out = (float2 *)pm->grad.ptr;
    float *rho = phi_q ? phi_q : (float *)pm->paint.ptr;
    HYMD_TRY(sort_particles(pm, x, nullptr, charges, n));
    HYMD_TRY(paint_sorted(pm, 1, nullptr, rho, nullptr));
    PmeK p;
    p.sigma2 = sigma * sigma;
    p.inv_g = 1.0 / (double)m.g;
    p.four_pi_elec = 4.0 * M_PI * elec_factor;
    KGeom g = kgeom(pm);
    double *partials = energy_q ? (double *)pm->partials.ptr : nullptr;
    int nparts = pm->kblocks;
    const XlineShape sh = xline_grad_shape(pm->fft, 1, pm->xline_w);
    if (sh.W > 0) {
        // z + y forward passes, then one x-line pass: the forward x
        // transform, the k-space pass and the inverse x transforms of
        // -i k_x psi and psi
        HYMD_TRY(fft_r2c_zy(pm->fft, rho, spec, 1, s));
        FieldK fk{};
        fk.T = 1;
        HYMD_TRY(launch_xline_grad(pm, spec, out, g, fk, sh, partials, &nparts, &p));
    } else {
        HYMD_TRY(fft_r2c_real(pm->fft, rho, spec, 1, s));
        pme_kspace_kernel<<<pm->kblocks, kBlock, 0, s>>>(spec, out, psi ? 1 : 0, 1, g, p, partials);
        HYMD_HIP(hipGetLastError());
        // x-inverse of -i k_x psi and psi
        HYMD_TRY(fft_inverse_x(pm->fft, out, 2, s));
    }
    float *field = elec_field;
    if (!field) {
        HYMD_TRY(pm->rscratch.reserve(sizeof(float) * 3 * m.g));
        field = (float *)pm->rscratch.ptr;
    }
    // the plane pass (or the y and z line passes) forms E_y, E_z from psi
    if (pm->fft.lds_plane) {
        HYMD_TRY(fft_c2r_yz_grad(pm->fft, out, field, 1, g.kfac[1], g.kfac[2], 1.f, s));
    } else {
        HYMD_TRY(pm->cscratch.reserve(sizeof(float2) * m.gc * 3));
        HYMD_TRY(fft_c2r_lines_grad(pm->fft, out, (float2 *)pm->cscratch.ptr, field, 1, g.kfac[1], g.kfac[2], 1.f,
                                    s));
    }
    if (psi) HYMD_TRY(fft_c2r_yz(pm->fft, out + m.gc, psi, 1, 1.f, s));
    HYMD_TRY(readout_sorted(pm, 3, 1, field, elec_forces, true, accumulate));
    if (energy_q) {
        double e = 0.0;
        HYMD_TRY(read_result(pm, nparts, 1, &e));
        *energy_q = e * (m.box[0] * m.box[1] * m.box[2]) - self_energy;
    }
    return HYMD_OK;
}

extern "C" int hymd_field_energy(hymd_pm *pm, const float *phi, int32_t n_types, const double *coupling,
                                 const double *shift, double constant, double *energy) {
    HYMD_REQUIRE(pm && phi && coupling && energy, "null argument");
    HYMD_REQUIRE(n_types >= 1 && n_types <= HYMD_MAX_TYPES, "n_types out of range");
    const Mesh &m = pm->mesh;
    FieldK p = field_params(n_types, coupling, shift, constant, 0.0, m);
    const int nb = pm->kblocks;
    double *part = (double *)pm->partials.ptr;
    if (n_types <= 4)
        field_energy_kernel<4><<<nb, kBlock, 0, pm->stream>>>(phi, m.g, p, part);
    else if (n_types <= 8)
        field_energy_kernel<8><<<nb, kBlock, 0, pm->stream>>>(phi, m.g, p, part);
    else
        field_energy_kernel<16><<<nb, kBlock, 0, pm->stream>>>(phi, m.g, p, part);
    HYMD_HIP(hipGetLastError());
    double e = 0.0;
    HYMD_TRY(read_result(pm, nb, 1, &e));
    *energy = e * cell_volume(pm);
    return HYMD_OK;
}

extern "C" int hymd_field_q_energy(hymd_pm *pm, const float *phi_q, const float *psi, double self_energy,
                                   double *energy) {
    HYMD_REQUIRE(pm && phi_q && psi && energy, "null argument");
    const Mesh &m = pm->mesh;
    const int nb = pm->kblocks;
    field_q_energy_kernel<<<nb, kBlock, 0, pm->stream>>>(phi_q, psi, m.g, (double *)pm->partials.ptr);
    HYMD_HIP(hipGetLastError());
    double e = 0.0;
    HYMD_TRY(read_result(pm, nb, 1, &e));
    *energy = e * cell_volume(pm) - self_energy;
    return HYMD_OK;
}

extern "C" int hymd_self_energy_q(const float *charges, int64_t n, double sigma, double elec_factor, double *out,
                                  void *stream) {
    HYMD_REQUIRE(out && (charges || n == 0) && n >= 0 && sigma > 0.0, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    const int nb = grid_for(n, kBlock, 1024);
    double *part = sync_scratch(nb + 1);  // nb partials, then the sum
    if (!part) return HYMD_ERR_NOMEM;
    self_energy_kernel<<<nb, kBlock, 0, s>>>(charges, n, part);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials(part, nb, 1, part + nb, s));
    double q2 = 0.0;
    HYMD_HIP(hipMemcpyAsync(&q2, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    *out = elec_factor * std::sqrt(1.0 / (2.0 * M_PI * sigma * sigma)) * q2;
    return HYMD_OK;
}

// ===========================================================================
// Variable-dielectric electrostatics: update_field_force_q_GPE
// (field.py:764-1112) and compute_field_energy_q_GPE (field.py:706-761)
// ===========================================================================
namespace hymd {

struct GpeK {
    int T;
    double eps[HYMD_MAX_TYPES];  // relative dielectric per type (config.dielectric_type)
    double q[HYMD_MAX_TYPES];    // charge per type (config.type_charges)
};

// den = sum_t phi_t, eps = sum_t eps_t phi_t / den where den > 1e-6 (the
// caller's previous value elsewhere, as np.divide(..., where=) leaves it);
// then phi_q /= eps where eps > 1e-6 (field.py:982-992)
__global__ void __launch_bounds__(kBlock) gpe_eps_kernel(const float *__restrict__ phi, int64_t g, GpeK p,
                                                          float *__restrict__ eps, float *__restrict__ phi_q,
                                                          float *__restrict__ den_out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g; c += stride) {
        double den = 0.0, num = 0.0;
        for (int t = 0; t < p.T; ++t) {
            const double v = phi[(int64_t)t * g + c];
            num += p.eps[t] * v;
            den += v;
        }
        float e = eps[c];
        if (den > 1e-6) e = (float)(num / den);
        eps[c] = e;
        den_out[c] = (float)den;
        if (e > 1e-6f) phi_q[c] = phi_q[c] / e;
    }
}

// The GPE's k-space multiplies of unnormalised spectra (scale = 1/G folds the
// r2c normalisation in; |k|^2 := 1 at k = 0, k_d = 0 on Nyquist planes):
//   0: out[d] = i k_d in            (gradient of the dielectric)
//   1: out[d] = -i k_d in / |k|^2   (field of a charge density, without 4 pi k_e)
//   2: out[d] = -i k_d in / |k|^2, out[3] = in / |k|^2   (field and potential)
//   3: out[3t + d] = -i k_d H in[t], t < T                 (force grids of V_t)
//   4: out[0] = H in                                       (the filter)
__global__ void __launch_bounds__(kBlock) gpe_kspace_kernel(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                             KGeom g, int mode, int T, double sigma2, double scale) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        if (kk == 0) k2 = 1.0;
        if (mode == 3) {
            const double h = exp(-0.5 * sigma2 * (kk == 0 ? 0.0 : k2)) * scale;
            for (int t = 0; t < T; ++t) {
                const float2 v = in[(int64_t)t * g.gc + kk];
                const double re = h * v.x, im = h * v.y;
                for (int d = 0; d < 3; ++d)
                    out[(int64_t)(3 * t + d) * g.gc + kk] = make_float2((float)(kd[d] * im), (float)(-kd[d] * re));
            }
            continue;
        }
        const float2 v = in[kk];
        if (mode == 4) {
            const double h = exp(-0.5 * sigma2 * (kk == 0 ? 0.0 : k2)) * scale;
            out[kk] = make_float2((float)(h * v.x), (float)(h * v.y));
            continue;
        }
        const double s = (mode == 0) ? scale : scale / k2;
        const double re = s * v.x, im = s * v.y;
        for (int d = 0; d < 3; ++d) {
            if (mode == 0)  // i k v = (-k v.y, k v.x)
                out[(int64_t)d * g.gc + kk] = make_float2((float)(-kd[d] * im), (float)(kd[d] * re));
            else            // -i k v = (k v.y, -k v.x)
                out[(int64_t)d * g.gc + kk] = make_float2((float)(kd[d] * im), (float)(-kd[d] * re));
        }
        if (mode == 2) out[(int64_t)3 * g.gc + kk] = make_float2((float)re, (float)im);
    }
}

// phi_eta_d /= eps where eps > 1e-6 (field.py:1007)
__global__ void __launch_bounds__(kBlock) gpe_eta_kernel(float *__restrict__ eta, const float *__restrict__ eps,
                                                          int64_t g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g; c += stride) {
        const float e = eps[c];
        if (e > 1e-6f)
            for (int d = 0; d < 3; ++d) eta[(int64_t)d * g + c] = eta[(int64_t)d * g + c] / e;
    }
}

// out = a (phi_q + pol)
__global__ void __launch_bounds__(kBlock) gpe_sum_kernel(const float *__restrict__ phi_q, const float *__restrict__ pol,
                                                          float a, float *__restrict__ out, int64_t g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g; c += stride)
        out[c] = a * (phi_q[c] + pol[c]);
}

// One polarisation update (field.py:1026-1040): pol <- w (-eta . E) +
// (1 - w) pol, per block the sum, sum of squares and maximum of |change|.
__global__ void __launch_bounds__(kBlock) gpe_pol_kernel(const float *__restrict__ eta, const float *__restrict__ E,
                                                          float *__restrict__ pol, int64_t g, float w,
                                                          double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    double mx = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g; c += stride) {
        float p = -(eta[c] * E[c] + eta[g + c] * E[g + c] + eta[2 * g + c] * E[2 * g + c]);
        const float prev = pol[c];
        p = w * p + (1.0f - w) * prev;
        const double diff = fabs((double)(p - prev));
        acc[0] += diff;
        acc[1] += diff * diff;
        mx = fmax(mx, diff);
        pol[c] = p;
    }
    block_reduce_store<2>(acc, partials + 3 * (int64_t)blockIdx.x);
    // the block maximum: wave shuffles, then the waves' maxima in LDS
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_down(mx, o, 64));
    __shared__ double wmax[kBlock / 64];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) m = fmax(m, wmax[k]);
        partials[3 * (int64_t)blockIdx.x + 2] = m;
    }
}

// elec_dot = sum_d E_d^2, contrib = elec_dot / den where den > 1e-6 (kept
// elsewhere), V_t = q_t psi - inv2e (eps_t - eps) contrib (field.py:1068-1086);
// per block sum eps elec_dot (the energy, field.py:757-759)
__global__ void __launch_bounds__(kBlock) gpe_vbar_kernel(const float *__restrict__ E, const float *__restrict__ psi,
                                                           const float *__restrict__ eps, const float *__restrict__ den,
                                                           GpeK p, float inv2e, int64_t g,
                                                           float *__restrict__ contrib, float *__restrict__ elec_dot,
                                                           float *__restrict__ vbar, double *__restrict__ partials) {
    double acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < g; c += stride) {
        const float e0 = E[c], e1 = E[g + c], e2 = E[2 * g + c];
        const float dot = e0 * e0 + e1 * e1 + e2 * e2;
        if (elec_dot) elec_dot[c] = dot;
        const float d = den[c];
        float cb = contrib[c];
        if (d > 1e-6f) cb = dot / d;
        contrib[c] = cb;
        const float ec = eps[c], ps = psi[c];
        for (int t = 0; t < p.T; ++t)
            vbar[(int64_t)t * g + c] = (float)p.q[t] * ps - inv2e * ((float)p.eps[t] - ec) * cb;
        acc[0] += (double)ec * (double)dot;
    }
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

}  // namespace hymd

extern "C" int hymd_update_field_force_q_gpe(hymd_pm *pm, const float *x, const int32_t *types, const float *charges,
                                             int64_t n, int32_t n_types, const float *phi,
                                             const double *dielectric_type, const double *type_charges, double sigma,
                                             double coulomb_constant, double conv_crit, double pol_mixing,
                                             int32_t conv_type, int32_t max_iter, float *phi_eps,
                                             const float *phi_pol_prev, float *phi_pol, float *elec_field_contrib,
                                             float *psi, float *elec_dot, float *vbar, float *elec_forces,
                                             double *energy, int32_t *iterations, double *delta_out) {
    HYMD_REQUIRE(pm && phi && dielectric_type && type_charges && phi_eps && elec_field_contrib && elec_forces,
                 "null argument");
    HYMD_REQUIRE(n >= 0 && (n == 0 || (x && types && charges)), "bad particle arrays");
    HYMD_REQUIRE(n_types >= 1 && n_types <= pm->max_types, "n_types out of range");
    HYMD_REQUIRE(conv_type >= 0 && conv_type <= 2, "conv_type must be 0 (max), 1 (sum) or 2 (sum of squares)");
    HYMD_REQUIRE(max_iter >= 0 && sigma >= 0.0 && coulomb_constant > 0.0, "bad parameters");
    const Mesh &m = pm->mesh;
    const int T = n_types;
    const int64_t G = m.g;
    hipStream_t s = pm->stream;
    GpeK p{};
    p.T = T;
    for (int t = 0; t < T; ++t) {
        p.eps[t] = dielectric_type[t];
        p.q[t] = type_charges[t];
    }
    // real scratch: phi_q, den, eta (3), E (3), psi, tmp, pol, V (T), force grids (3T)
    HYMD_TRY(pm->rscratch.reserve(sizeof(float) * (size_t)G * (11 + 4 * (size_t)T)));
    float *phq = (float *)pm->rscratch.ptr, *den = phq + G, *eta = den + G, *E = eta + 3 * G, *psi_i = E + 3 * G;
    float *tmp = psi_i + G, *pol_i = tmp + G, *vb_i = pol_i + G, *fm = vb_i + (int64_t)T * G;
    float *pol = phi_pol ? phi_pol : pol_i;
    float *vb = vbar ? vbar : vb_i;
    float2 *spec = (float2 *)pm->spec.ptr, *grad = (float2 *)pm->grad.ptr;
    const KGeom g = kgeom(pm);
    const double inv_g = 1.0 / (double)G;
    const double s2 = sigma * sigma;
    const int kb = pm->kblocks;
    const int eb = grid_for(G, kBlock, 4096);
    HYMD_TRY(pm->partials.reserve(sizeof(double) * 3 * (size_t)eb));
    double *part = (double *)pm->partials.ptr;
    // 1. smeared charge density (field.py:971-980)
    HYMD_TRY(sort_particles(pm, x, nullptr, charges, n));
    HYMD_TRY(paint_sorted(pm, 1, nullptr, phq, nullptr));
    HYMD_TRY(fft_r2c_real(pm->fft, phq, spec, 1, s));
    gpe_kspace_kernel<<<kb, kBlock, 0, s>>>(spec, grad, g, 4, 1, s2, inv_g);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(fft_c2r(pm->fft, grad, phq, 1, 1.f, s));
    // 2-3. dielectric, phi_q / eps
    gpe_eps_kernel<<<eb, kBlock, 0, s>>>(phi, G, p, phi_eps, phq, den);
    HYMD_HIP(hipGetLastError());
    // 4. grad(eps) / eps
    HYMD_TRY(fft_r2c_real(pm->fft, phi_eps, spec, 1, s));
    gpe_kspace_kernel<<<kb, kBlock, 0, s>>>(spec, grad, g, 0, 1, s2, inv_g);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(fft_c2r(pm->fft, grad, eta, 3, 1.f, s));
    gpe_eta_kernel<<<eb, kBlock, 0, s>>>(eta, phi_eps, G);
    HYMD_HIP(hipGetLastError());
    // 5. polarisation charges from the caller's previous grid
    if (phi_pol_prev)
        HYMD_HIP(hipMemcpyAsync(pol, phi_pol_prev, sizeof(float) * G, hipMemcpyDeviceToDevice, s));
    else
        HYMD_TRY(zero_words(pol, G, s));
    std::vector<double> hp(3 * (size_t)eb);
    int it = 0;
    double delta = 1.0;
    while (it < max_iter && delta > conv_crit) {
        gpe_sum_kernel<<<eb, kBlock, 0, s>>>(phq, pol, 1.f, tmp, G);
        HYMD_TRY(fft_r2c_real(pm->fft, tmp, spec, 1, s));
        gpe_kspace_kernel<<<kb, kBlock, 0, s>>>(spec, grad, g, 1, 1, s2, inv_g);
        HYMD_HIP(hipGetLastError());
        HYMD_TRY(fft_c2r(pm->fft, grad, E, 3, 1.f, s));
        gpe_pol_kernel<<<eb, kBlock, 0, s>>>(eta, E, pol, G, (float)pol_mixing, part);
        HYMD_HIP(hipGetLastError());
        HYMD_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * 3 * eb, hipMemcpyDeviceToHost, s));
        HYMD_HIP(hipStreamSynchronize(s));
        double sum = 0.0, sq = 0.0, mx = 0.0;
        for (int b = 0; b < eb; ++b) {
            sum += hp[3 * b];
            sq += hp[3 * b + 1];
            mx = std::max(mx, hp[3 * b + 2]);
        }
        delta = conv_type == 0 ? mx : (conv_type == 1 ? sum : sq);
        ++it;
    }
    // 6. potential and field: psi = 4 pi k_e (phi_q + pol) / |k|^2
    const double eps0_inv = 4.0 * M_PI * coulomb_constant;
    gpe_sum_kernel<<<eb, kBlock, 0, s>>>(phq, pol, (float)eps0_inv, tmp, G);
    HYMD_TRY(fft_r2c_real(pm->fft, tmp, spec, 1, s));
    gpe_kspace_kernel<<<kb, kBlock, 0, s>>>(spec, grad, g, 2, 1, s2, inv_g);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(fft_c2r(pm->fft, grad, E, 4, 1.f, s));  // E (3 grids), then psi_i
    if (psi) HYMD_HIP(hipMemcpyAsync(psi, psi_i, sizeof(float) * G, hipMemcpyDeviceToDevice, s));
    // 7. |E|^2, contrib, V_t and the energy partials
    gpe_vbar_kernel<<<eb, kBlock, 0, s>>>(E, psi_i, phi_eps, den, p, (float)(0.5 / eps0_inv), G, elec_field_contrib,
                                          elec_dot, vb, part);
    HYMD_HIP(hipGetLastError());
    double e_sum = 0.0;
    if (energy) {
        HYMD_HIP(hipMemcpyAsync(hp.data(), part, sizeof(double) * eb, hipMemcpyDeviceToHost, s));
        HYMD_HIP(hipStreamSynchronize(s));
        for (int b = 0; b < eb; ++b) e_sum += hp[b];
    }
    // 8. force grids -i k_d H V_t and the per-type readout
    HYMD_TRY(fft_r2c_real(pm->fft, vb, spec, T, s));
    gpe_kspace_kernel<<<kb, kBlock, 0, s>>>(spec, grad, g, 3, T, s2, inv_g);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(fft_c2r(pm->fft, grad, fm, 3 * T, 1.f, s));
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
    HYMD_TRY(readout_sorted(pm, 3, T, fm, elec_forces, false));
    if (energy) *energy = cell_volume(pm) * (0.5 / eps0_inv) * e_sum;
    if (iterations) *iterations = it;
    if (delta_out) *delta_out = delta;
    return HYMD_OK;
}
