// Batched replicas: many independent systems, each with its own particle
// range, mesh, box and filter width, binned, painted, transformed and reduced
// to spectral moments (P, Q, phi~(0)) by one sequence of launches that can be
// captured once and replayed as a hipGraph.  From the moments the host forms
// every replica's energy and its chi / kappa / sigma derivatives.
//
// The per-replica arithmetic is the single-mesh path's (field_dev.h device
// functions, the same FFT plans, the same moments block count and reduction),
// so every replica's moments equal those of hymd_field_moments bit for bit.
//
// Slots: replicas are laid out internally grouped by mesh shape (stable), so
// that each group's grids are contiguous and one batched FFT covers a group.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "fft.h"
#include "field_dev.h"

namespace hymd {

struct RepGeom {
    int mx, my, mz, zh;
    int row_off;       // first global mesh row of the replica
    int nb;            // moments blocks
    int64_t g, gc;
    int64_t grid_off;  // float offset of its T real grids
    int64_t spec_off;  // complex offset of its T spectra
    int64_t part_off;  // first partial row
    int64_t grad_off;  // complex offset of its 3T gradient spectra
    int64_t fgrid_off; // float offset of its 3T force grids
    double inv_h[3];
    double kfac[3];
    double out_scale;  // 2^-32 / dV
};

struct RepGroup {
    int mesh[3];
    int first, count;  // slots
    int nb;
    FftPlan plan;
};

// particle i -> caller replica r (poff ascending, R + 1 entries)
__device__ __forceinline__ int find_range(const int64_t *__restrict__ off, int R, int64_t i) {
    int lo = 0, hi = R;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Particles are consecutive per replica, so a wavefront's 64 particles
// nearly always belong to one replica: one uniform load of the chunk table
// (chunk_in[c] = (slot, end of its replica) of particle 64c, built with the
// handle) gives it, and only the lanes past its end (a wavefront that
// straddles two replicas) search on their own.
__global__ void __launch_bounds__(kBlock) batch_key_kernel(const float *__restrict__ x, int64_t n,
                                                            const int64_t *__restrict__ poff, int R,
                                                            const int *__restrict__ slot_of,
                                                            const int2 *__restrict__ chunk_in,
                                                            const RepGeom *__restrict__ geo, int2 *__restrict__ keys,
                                                            int *__restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x; i0 < n;
         i0 += stride) {  // uniform per wavefront
        const int64_t i = i0 + threadIdx.x;
        const int wbase = __builtin_amdgcn_readfirstlane((int)(i0 + (threadIdx.x & ~63u)));
        int key = -1;
        if (wbase < n) {
            const int2 ck = chunk_in[wbase >> 6];
            const RepGeom *G = geo + ck.x;
            if (i < n && i >= ck.y) G = geo + slot_of[find_range(poff, R, i)];
            if (i < n) {
                const CicAxis cx = cic_axis((double)x[3 * i + 0], G->inv_h[0], G->mx);
                const CicAxis cy = cic_axis((double)x[3 * i + 1], G->inv_h[1], G->my);
                key = G->row_off + cx.i0 * G->my + cy.i0;
            }
        }
        const int rank = rank_in_bin_all(key, counts);
        if (i < n) keys[i] = make_int2(key, rank);
    }
}

// One wavefront per global mesh row (all replicas), kBatchRows rows per
// workgroup: the row paint of the single-mesh path with the replica's
// geometry from the table.  A wavefront issues the record loads of up to four
// visits per lane before their deposits, so each row waits on memory about
// twice, not once per visit.
constexpr int kBatchRows = kBlock / 64;

__global__ void __launch_bounds__(kBlock) batch_paint_rows_kernel(const float4 *__restrict__ rec,
                                                                   const int *__restrict__ start,
                                                                   const RepGeom *__restrict__ geo,
                                                                   const int *__restrict__ row_slot, int rows,
                                                                   int max_mz, PaintArgs pa,
                                                                   float *__restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ unsigned long long acc_all[];  // kBatchRows x T x max_mz
    const int T = pa.n_types;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grow = (int)xcd_block(blockIdx.x, gridDim.x) * kBatchRows + wave;
    const bool active = grow < rows;
    unsigned long long *acc = acc_all + (int64_t)wave * T * max_mz;
    const RepGeom *G = geo + row_slot[active ? grow : rows - 1];
    const int row = grow - G->row_off;
    const int ix = row / G->my, iy = row - ix * G->my;
    const int mz = G->mz;
    for (int i = lane; i < T * mz; i += 64) acc[i] = 0ull;
    int rs[4], rn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = q >> 1, b = q & 1;
        const int bx = a ? (ix == 0 ? G->mx - 1 : ix - 1) : ix;
        const int by = b ? (iy == 0 ? G->my - 1 : iy - 1) : iy;
        const int br = G->row_off + bx * G->my + by;
        rs[q] = active ? start[br] : 0;
        rn[q] = active ? start[br + 1] - rs[q] : 0;
    }
    __syncthreads();
    const int tot = rn[0] + rn[1] + rn[2] + rn[3];
    for (int v0 = 0; v0 < tot; v0 += 4 * 64) {
        float4 r[4];
        int qq[4];
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
                r[u] = rec[rs[q] + off];
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (qq[u] < 0) continue;
            const int t = __float_as_int(r[u].w);
            if ((unsigned)t >= (unsigned)T) continue;
            row_deposit(r[u], qq[u] >> 1, qq[u] & 1, pa.mass[t], pa.scale, G->inv_h, G->mx, G->my, mz,
                        acc + t * mz);
        }
    }
    __syncthreads();
    if (!active) return;
    float *o = out + G->grid_off + (int64_t)row * mz;
    for (int i = lane; i < T * mz; i += 64) {
        const int t = i / mz, z = i - t * mz;
        o[(int64_t)t * G->g + z] = (float)((double)(long long)acc[i] * G->out_scale);
    }
}

// grid (k-space blocks, slots of the group): the gradient spectra
// -i k_d V_t of each replica, V_t = H sum_s A_ts H rho_s / G, with the
// arithmetic of the single-mesh k-space pass (field.hip), so each replica's
// force grids equal hymd_field_moments' bit for bit.  coupling: per slot T x T.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) batch_kspace_kernel(const float2 *__restrict__ spec,
                                                               const RepGeom *__restrict__ geo, int slot0, int T,
                                                               const double *__restrict__ sigma2,
                                                               const double *__restrict__ coupling,
                                                               float2 *__restrict__ grad) {
    const int slot = slot0 + (int)blockIdx.y;
    const RepGeom &G = geo[slot];
    KGeom g;
    g.mx = G.mx;
    g.my = G.my;
    g.mz = G.mz;
    g.zh = G.zh;
    g.gc = G.gc;
    for (int d = 0; d < 3; ++d) g.kfac[d] = G.kfac[d];
    const double s2 = sigma2[slot], inv_g = 1.0 / (double)G.g;
    const double *A = coupling + (int64_t)slot * T * T;
    const float2 *sp = spec + G.spec_off;
    float2 *gr = grad + G.grad_off;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
        double k[3], kd[3], wgt;
        kvec(g, kk, k, kd, wgt);
        const double k2 = k[0] * k[0] + k[1] * k[1] + k[2] * k[2];
        const float H = (float)exp(-0.5 * s2 * k2);
        const float hn = (float)((double)H * inv_g);
        float2 ph[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t)
            if (t < T) ph[t] = cscale(sp[(int64_t)t * g.gc + kk], hn);
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= T) continue;
            float2 v = make_float2(0.f, 0.f);
#pragma unroll
            for (int u = 0; u < MAXT; ++u)
                if (u < T) {
                    const float a = (float)A[t * T + u];
                    v.x += a * ph[u].x;
                    v.y += a * ph[u].y;
                }
            v = cscale(v, H);
            float2 *gt = gr + (int64_t)(3 * t) * g.gc + kk;
            gt[0] = make_float2((float)kd[0] * v.y, -(float)kd[0] * v.x);
            gt[g.gc] = make_float2((float)kd[1] * v.y, -(float)kd[1] * v.x);
            gt[2 * g.gc] = make_float2((float)kd[2] * v.y, -(float)kd[2] * v.x);
        }
    }
}

// Forces of the binned particles: each reads its replica's 3 force grids of
// its type (the single-mesh readout's arithmetic), written to row perm[p].
// Every particle is binned and slots hold consecutive global rows, so slot s
// holds binned particles [pstart[s], pstart[s + 1]) whatever the positions:
// the chunk table (chunk_binned[c] = (slot, pstart[slot + 1]) of binned
// particle 64c, built with the handle) gives a wavefront's slot in one
// uniform load, and lanes past its end step to the next slots.
__global__ void __launch_bounds__(kBlock) batch_readout_kernel(const float *__restrict__ fgrid,
                                                                const float4 *__restrict__ rec,
                                                                const int *__restrict__ perm,
                                                                const int2 *__restrict__ chunk_binned,
                                                                const int *__restrict__ pstart,
                                                                const RepGeom *__restrict__ geo, int T,
                                                                int64_t n, float *__restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x; p0 < n; p0 += stride) {
        const int64_t p = p0 + threadIdx.x;
        const int wbase = __builtin_amdgcn_readfirstlane((int)(p0 + (threadIdx.x & ~63u)));
        if (wbase >= n) continue;
        const int2 ck = chunk_binned[wbase >> 6];
        if (p >= n) continue;
        int sl = ck.x;
        if (p >= ck.y)
            while (p >= pstart[sl + 1]) ++sl;
        const RepGeom &G = geo[sl];
        const float4 r = rec[p];
        const int t = __float_as_int(r.w);
        const int64_t o = perm[p];
        if ((unsigned)t >= (unsigned)T) {
            out[3 * o] = out[3 * o + 1] = out[3 * o + 2] = 0.f;
            continue;
        }
        const CicAxis cx = cic_axis((double)r.x, G.inv_h[0], G.mx);
        const CicAxis cy = cic_axis((double)r.y, G.inv_h[1], G.my);
        const CicAxis cz = cic_axis((double)r.z, G.inv_h[2], G.mz);
        const float *gt = fgrid + G.fgrid_off + (int64_t)t * 3 * G.g;
        float v[8][3];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int a = c >> 2, b = (c >> 1) & 1, d = c & 1;
            const int64_t idx =
                ((int64_t)(a ? cx.i1 : cx.i0) * G.my + (b ? cy.i1 : cy.i0)) * G.mz + (d ? cz.i1 : cz.i0);
#pragma unroll
            for (int q = 0; q < 3; ++q) v[c][q] = gt[q * G.g + idx];
        }
        double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int a = c >> 2, b = (c >> 1) & 1, d = c & 1;
            const double w = ((a ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0)) * (d ? cz.w1 : cz.w0);
#pragma unroll
            for (int q = 0; q < 3; ++q) acc[q] += w * (double)v[c][q];
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) out[3 * o + q] = (float)acc[q];
    }
}

// grid (nb of the group, slots of the group)
template <int MAXT>
__global__ void __launch_bounds__(kBlock) batch_moments_kernel(const float2 *__restrict__ spec,
                                                                const RepGeom *__restrict__ geo, int slot0, int T,
                                                                const double *__restrict__ sigma2,
                                                                double *__restrict__ partials) {
    constexpr int ROW = 2 * (MAXT * (MAXT + 1) / 2) + MAXT;
    const int slot = slot0 + (int)blockIdx.y;
    const RepGeom &G = geo[slot];
    KGeom g;
    g.mx = G.mx;
    g.my = G.my;
    g.mz = G.mz;
    g.zh = G.zh;
    g.gc = G.gc;
    for (int d = 0; d < 3; ++d) g.kfac[d] = G.kfac[d];
    moments_block<MAXT>(spec + G.spec_off, g, T, sigma2[slot], 1.0 / (double)G.g, blockIdx.x, gridDim.x,
                        partials + (G.part_off + blockIdx.x) * ROW);
}

}  // namespace hymd

using namespace hymd;

struct hymd_batch {
    int R = 0, T = 0, rows = 0, row_len = 0;
    int64_t n = 0;
    hipStream_t stream = nullptr;  // the library's own non-blocking stream (capturable)
    hipEvent_t ev_in = nullptr;
    PaintArgs pa;
    std::vector<RepGeom> geo;           // per slot
    std::vector<int> slot_of;           // caller replica -> slot
    std::vector<int64_t> poff;          // caller particle offsets (R + 1)
    std::vector<std::unique_ptr<RepGroup>> groups;
    DevBuf d_geo, d_slot_of, d_poff, d_sigma2, d_nb, d_row0;
    DevBuf d_chunk_in, d_chunk_binned, d_pstart, d_row_slot;  // lookup tables (see the kernels)
    DevBuf keys, perm, xs, counts, start, tiles;  // xs: binned records (x, y, z, type bits)
    DevBuf rho, spec, partials, result;
    DevBuf grad, fgrid, d_coupling;  // force stages: 3T spectra and grids per replica, T x T coupling per slot
    HostBuf h_sigma2, h_result, h_coupling;
    int64_t grad_total = 0, fgrid_total = 0;
    int max_mz = 0;
    // captured sequences: [0] moments, [1] moments + forces
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    const float *cap_x[2] = {nullptr, nullptr};
    const int32_t *cap_types[2] = {nullptr, nullptr};
    float *cap_forces = nullptr;
    ~hymd_batch() {
        for (int i = 0; i < 2; ++i) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
        }
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

// forces (device (n,3)) or NULL: the moments only
static int batch_enqueue(hymd_batch *b, const float *x, const int32_t *types, float *forces) {
    hipStream_t s = b->stream;
    const int S = b->R;
    if (b->n > 0)
        batch_key_kernel<<<grid_for(b->n), kBlock, 0, s>>>(x, b->n, (const int64_t *)b->d_poff.ptr, b->R,
                                                           (const int *)b->d_slot_of.ptr,
                                                           (const int2 *)b->d_chunk_in.ptr,
                                                           (const RepGeom *)b->d_geo.ptr, (int2 *)b->keys.ptr,
                                                           (int *)b->counts.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan((int *)b->counts.ptr, (int *)b->start.ptr, b->rows, (int *)b->tiles.ptr, nullptr,
                            true, s));
    HYMD_TRY(launch_key_scatter(x, types, nullptr, b->n, (const int2 *)b->keys.ptr, (const int *)b->start.ptr,
                                (int *)b->perm.ptr, (float4 *)b->xs.ptr, nullptr, s));
    const int pblocks = (b->rows + kBatchRows - 1) / kBatchRows;
    batch_paint_rows_kernel<<<pblocks, kBlock, sizeof(long long) * kBatchRows * b->T * b->max_mz, s>>>(
        (const float4 *)b->xs.ptr, (const int *)b->start.ptr, (const RepGeom *)b->d_geo.ptr,
        (const int *)b->d_row_slot.ptr, b->rows, b->max_mz,
        b->pa, (float *)b->rho.ptr);
    HYMD_HIP(hipGetLastError());
    for (const auto &gp : b->groups) {
        const RepGeom &G0 = b->geo[gp->first];
        HYMD_TRY(fft_r2c_real(gp->plan, (const float *)b->rho.ptr + G0.grid_off,
                              (float2 *)b->spec.ptr + G0.spec_off, gp->count * b->T, s));
        dim3 grid((unsigned)gp->nb, (unsigned)gp->count);
        if (b->T <= 2)
            batch_moments_kernel<2><<<grid, kBlock, 0, s>>>((const float2 *)b->spec.ptr,
                                                            (const RepGeom *)b->d_geo.ptr, gp->first, b->T,
                                                            (const double *)b->d_sigma2.ptr,
                                                            (double *)b->partials.ptr);
        else if (b->T <= 4)
            batch_moments_kernel<4><<<grid, kBlock, 0, s>>>((const float2 *)b->spec.ptr,
                                                            (const RepGeom *)b->d_geo.ptr, gp->first, b->T,
                                                            (const double *)b->d_sigma2.ptr,
                                                            (double *)b->partials.ptr);
        else
            batch_moments_kernel<8><<<grid, kBlock, 0, s>>>((const float2 *)b->spec.ptr,
                                                            (const RepGeom *)b->d_geo.ptr, gp->first, b->T,
                                                            (const double *)b->d_sigma2.ptr,
                                                            (double *)b->partials.ptr);
        HYMD_HIP(hipGetLastError());
        if (forces) {
            dim3 kgrid((unsigned)std::min(grid_for(G0.gc), 1024), (unsigned)gp->count);
            const double *cp = (const double *)b->d_coupling.ptr;
            float2 *grad = (float2 *)b->grad.ptr;
            if (b->T <= 4)
                batch_kspace_kernel<4><<<kgrid, kBlock, 0, s>>>((const float2 *)b->spec.ptr,
                                                                (const RepGeom *)b->d_geo.ptr, gp->first, b->T,
                                                                (const double *)b->d_sigma2.ptr, cp, grad);
            else
                batch_kspace_kernel<8><<<kgrid, kBlock, 0, s>>>((const float2 *)b->spec.ptr,
                                                                (const RepGeom *)b->d_geo.ptr, gp->first, b->T,
                                                                (const double *)b->d_sigma2.ptr, cp, grad);
            HYMD_HIP(hipGetLastError());
            HYMD_TRY(fft_c2r(gp->plan, grad + G0.grad_off, (float *)b->fgrid.ptr + G0.fgrid_off,
                             gp->count * 3 * b->T, 1.f, s));
        }
    }
    if (forces && b->n > 0) {
        batch_readout_kernel<<<grid_for(b->n), kBlock, 0, s>>>((const float *)b->fgrid.ptr, (const float4 *)b->xs.ptr,
                                                              (const int *)b->perm.ptr,
                                                              (const int2 *)b->d_chunk_binned.ptr,
                                                              (const int *)b->d_pstart.ptr,
                                                              (const RepGeom *)b->d_geo.ptr, b->T, b->n, forces);
        HYMD_HIP(hipGetLastError());
    }
    HYMD_TRY(reduce_partials_batched((const double *)b->partials.ptr, (const int *)b->d_nb.ptr,
                                     (const int64_t *)b->d_row0.ptr, b->row_len, (double *)b->result.ptr, S, s));
    HYMD_HIP(hipMemcpyAsync(b->h_result.ptr, b->result.ptr, sizeof(double) * b->row_len * S,
                            hipMemcpyDeviceToHost, s));
    return HYMD_OK;
}

extern "C" int hymd_batch_create(hymd_batch **out, int32_t n_replicas, const int32_t *mesh, const double *box,
                                 const int64_t *offsets, int32_t n_types, const double *mass) {
    HYMD_REQUIRE(out && mesh && box && offsets, "null argument");
    HYMD_REQUIRE(n_replicas >= 1 && n_replicas <= 4096, "n_replicas must be in [1, 4096]");
    HYMD_REQUIRE(n_types >= 1 && n_types <= 8, "batched moments support n_types <= 8");
    const int R = n_replicas, T = n_types;
    HYMD_REQUIRE(offsets[0] == 0, "offsets[0] must be 0");
    for (int r = 0; r < R; ++r) {
        HYMD_REQUIRE(offsets[r + 1] >= offsets[r], "offsets must be non-decreasing");
        for (int d = 0; d < 3; ++d) {
            const int M = mesh[3 * r + d];
            HYMD_REQUIRE(M >= 4 && M <= 1024 && M % 2 == 0, "replica %d: mesh edge must be even and in [4, 1024]", r);
            HYMD_REQUIRE(box[3 * r + d] > 0.0 && std::isfinite(box[3 * r + d]), "replica %d: bad box", r);
        }
    }
    HYMD_REQUIRE(offsets[R] < (int64_t(1) << 31), "too many particles");
    std::unique_ptr<hymd_batch> b(new hymd_batch());
    b->R = R;
    b->T = T;
    b->n = offsets[R];
    HYMD_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    HYMD_HIP(hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming));
    b->poff.assign(offsets, offsets + R + 1);
    b->pa.n_types = T;
    b->pa.scale = std::ldexp(1.0, HYMD_FIXED_SHIFT);
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) b->pa.mass[t] = (mass && t < T) ? mass[t] : 1.0;
    b->row_len = moments_row_len(T);
    // slots: replicas grouped by mesh shape, in order of first appearance
    std::vector<int> order;
    std::vector<char> used(R, 0);
    for (int r = 0; r < R; ++r) {
        if (used[r]) continue;
        auto grp = std::unique_ptr<RepGroup>(new RepGroup());
        for (int d = 0; d < 3; ++d) grp->mesh[d] = mesh[3 * r + d];
        grp->first = (int)order.size();
        for (int q = r; q < R; ++q)
            if (!used[q] && std::equal(mesh + 3 * q, mesh + 3 * q + 3, grp->mesh)) {
                used[q] = 1;
                order.push_back(q);
            }
        grp->count = (int)order.size() - grp->first;
        b->groups.push_back(std::move(grp));
    }
    b->slot_of.assign(R, 0);
    b->geo.resize(R);
    int rows = 0;
    int64_t goff = 0, soff = 0, poff = 0;
    for (auto &gp : b->groups) {
        for (int k = 0; k < gp->count; ++k) {
            const int slot = gp->first + k, r = order[slot];
            b->slot_of[r] = slot;
            RepGeom &G = b->geo[slot];
            G.mx = gp->mesh[0];
            G.my = gp->mesh[1];
            G.mz = gp->mesh[2];
            G.zh = G.mz / 2 + 1;
            G.g = (int64_t)G.mx * G.my * G.mz;
            G.gc = (int64_t)G.mx * G.my * G.zh;
            double vol = 1.0;
            for (int d = 0; d < 3; ++d) {
                const double L = box[3 * r + d];
                G.inv_h[d] = (double)gp->mesh[d] / L;
                G.kfac[d] = 2.0 * M_PI / L;
                vol *= L;
            }
            G.out_scale = std::ldexp(1.0, -HYMD_FIXED_SHIFT) / (vol / (double)G.g);
            G.row_off = rows;
            G.grid_off = goff;
            G.spec_off = soff;
            G.nb = moments_blocks(G.gc);
            G.part_off = poff;
            G.grad_off = 3 * soff;
            G.fgrid_off = 3 * goff;
            rows += G.mx * G.my;
            goff += (int64_t)T * G.g;
            soff += (int64_t)T * G.gc;
            poff += G.nb;
        }
        gp->nb = b->geo[gp->first].nb;
        Mesh m;
        m.mx = gp->mesh[0];
        m.my = gp->mesh[1];
        m.mz = gp->mesh[2];
        m.zh = m.mz / 2 + 1;
        m.g = (int64_t)m.mx * m.my * m.mz;
        m.gc = (int64_t)m.mx * m.my * m.zh;
        for (int d = 0; d < 3; ++d) {
            m.box[d] = 1.0;
            m.inv_h[d] = (double)m.mx;
        }
        HYMD_TRY(fft_plan_init(gp->plan, m));
    }
    b->rows = rows;
    for (const RepGeom &G : b->geo) b->max_mz = std::max(b->max_mz, G.mz);
    HYMD_REQUIRE((int64_t)T * b->max_mz <= 2048, "n_types x the largest z edge must be <= 2048 (got %d x %d)", T,
                 b->max_mz);
    b->grad_total = 3 * soff;
    b->fgrid_total = 3 * goff;
    const size_t nn = (size_t)std::max<int64_t>(b->n, 1);
    std::vector<int> nbs(R);
    std::vector<int64_t> row0(R);
    for (int sl = 0; sl < R; ++sl) {
        nbs[sl] = b->geo[sl].nb;
        row0[sl] = b->geo[sl].part_off;
    }
    // lookup tables: per 64-particle chunk of the input order, the slot of its
    // first particle and the end of that particle's replica; per slot, its
    // first binned particle; per 64-particle chunk of the binned order, the
    // slot of its first particle and that slot's end; per global row, its slot
    const int64_t nchunks = (b->n + 63) / 64;
    std::vector<int2> chunk_in((size_t)std::max<int64_t>(nchunks, 1)), chunk_binned(chunk_in.size());
    std::vector<int> pstart(R + 1, 0), row_slot((size_t)std::max(rows, 1));
    for (int sl = 0; sl < R; ++sl) {
        const int r = order[sl];
        pstart[sl + 1] = pstart[sl] + (int)(b->poff[r + 1] - b->poff[r]);
        for (int q = 0; q < b->geo[sl].mx * b->geo[sl].my; ++q) row_slot[b->geo[sl].row_off + q] = sl;
    }
    for (int64_t c = 0, r = 0, sl = 0; c < nchunks; ++c) {
        const int64_t i = 64 * c;
        while (i >= b->poff[r + 1]) ++r;  // the replica holding particle i (skips empty ones)
        while (i >= pstart[sl + 1]) ++sl;
        chunk_in[c] = make_int2(b->slot_of[r], (int)b->poff[r + 1]);
        chunk_binned[c] = make_int2((int)sl, pstart[sl + 1]);
    }
    HYMD_TRY(b->d_chunk_in.reserve(sizeof(int2) * chunk_in.size()));
    HYMD_TRY(b->d_chunk_binned.reserve(sizeof(int2) * chunk_binned.size()));
    HYMD_TRY(b->d_pstart.reserve(sizeof(int) * (R + 1)));
    HYMD_TRY(b->d_row_slot.reserve(sizeof(int) * row_slot.size()));
    HYMD_TRY(b->d_geo.reserve(sizeof(RepGeom) * R));
    HYMD_TRY(b->d_slot_of.reserve(sizeof(int) * R));
    HYMD_TRY(b->d_poff.reserve(sizeof(int64_t) * (R + 1)));
    HYMD_TRY(b->d_sigma2.reserve(sizeof(double) * R));
    HYMD_TRY(b->d_nb.reserve(sizeof(int) * R));
    HYMD_TRY(b->d_row0.reserve(sizeof(int64_t) * R));
    HYMD_TRY(b->keys.reserve(sizeof(int2) * nn));
    HYMD_TRY(b->perm.reserve(sizeof(int) * nn));
    HYMD_TRY(b->xs.reserve(sizeof(float4) * nn));
    HYMD_TRY(b->counts.reserve(sizeof(int) * (rows + 2)));
    HYMD_TRY(b->start.reserve(sizeof(int) * (rows + 1)));
    HYMD_TRY(b->tiles.reserve(sizeof(int) * (scan_tiles(rows) + 1)));
    HYMD_TRY(b->rho.reserve(sizeof(float) * std::max<int64_t>(goff, 1)));
    HYMD_TRY(b->spec.reserve(sizeof(float2) * std::max<int64_t>(soff, 1)));
    HYMD_TRY(b->partials.reserve(sizeof(double) * b->row_len * std::max<int64_t>(poff, 1)));
    HYMD_TRY(b->result.reserve(sizeof(double) * b->row_len * R));
    HYMD_TRY(b->h_sigma2.reserve(sizeof(double) * R));
    HYMD_TRY(b->h_result.reserve(sizeof(double) * b->row_len * R));
    HYMD_TRY(b->h_coupling.reserve(sizeof(double) * T * T * R));
    HYMD_TRY(b->d_coupling.reserve(sizeof(double) * T * T * R));
    hipStream_t s = b->stream;
    HYMD_HIP(hipMemcpyAsync(b->d_geo.ptr, b->geo.data(), sizeof(RepGeom) * R, hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_slot_of.ptr, b->slot_of.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_poff.ptr, b->poff.data(), sizeof(int64_t) * (R + 1), hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_nb.ptr, nbs.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_chunk_in.ptr, chunk_in.data(), sizeof(int2) * chunk_in.size(), hipMemcpyHostToDevice,
                            s));
    HYMD_HIP(hipMemcpyAsync(b->d_chunk_binned.ptr, chunk_binned.data(), sizeof(int2) * chunk_binned.size(),
                            hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_pstart.ptr, pstart.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_row_slot.ptr, row_slot.data(), sizeof(int) * row_slot.size(), hipMemcpyHostToDevice,
                            s));
    HYMD_HIP(hipMemcpyAsync(b->d_row0.ptr, row0.data(), sizeof(int64_t) * R, hipMemcpyHostToDevice, s));
    HYMD_TRY(zero_words(b->counts.ptr, rows + 2, s));
    HYMD_HIP(hipStreamSynchronize(s));
    *out = b.release();
    return HYMD_OK;
}

extern "C" int hymd_batch_destroy(hymd_batch *b) {
    if (!b) return HYMD_OK;
    (void)hipStreamSynchronize(b->stream);
    delete b;
    return HYMD_OK;
}

static int batch_run(hymd_batch *b, const float *x, const int32_t *types, const double *sigma,
                     const double *coupling, double *moments, float *forces, int32_t use_graph, void *stream) {
    HYMD_REQUIRE(b && sigma && moments && ((x && types) || b->n == 0), "null argument");
    const int R = b->R, T = b->T;
    double *hs = (double *)b->h_sigma2.ptr;
    for (int r = 0; r < R; ++r) {
        HYMD_REQUIRE(sigma[r] >= 0.0 && std::isfinite(sigma[r]), "replica %d: sigma must be finite and >= 0", r);
        hs[b->slot_of[r]] = sigma[r] * sigma[r];
    }
    const int gi = forces ? 1 : 0;
    if (forces) {
        HYMD_REQUIRE(coupling, "forces need the coupling matrices");
        // the force stages' buffers are sized on first use
        HYMD_TRY(b->grad.reserve(sizeof(float2) * std::max<int64_t>(b->grad_total, 1)));
        HYMD_TRY(b->fgrid.reserve(sizeof(float) * std::max<int64_t>(b->fgrid_total, 1)));
        double *hc = (double *)b->h_coupling.ptr;
        for (int r = 0; r < R; ++r)
            std::memcpy(hc + (int64_t)b->slot_of[r] * T * T, coupling + (int64_t)r * T * T, sizeof(double) * T * T);
    }
    hipStream_t s = b->stream;
    // inputs written on the caller's stream are ready before the batch reads them
    HYMD_HIP(hipEventRecord(b->ev_in, (hipStream_t)stream));
    HYMD_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    HYMD_HIP(hipMemcpyAsync(b->d_sigma2.ptr, hs, sizeof(double) * R, hipMemcpyHostToDevice, s));
    if (forces)
        HYMD_HIP(hipMemcpyAsync(b->d_coupling.ptr, b->h_coupling.ptr, sizeof(double) * T * T * R,
                                hipMemcpyHostToDevice, s));
    if (use_graph) {
        if (!b->exec[gi] || b->cap_x[gi] != x || b->cap_types[gi] != types || (forces && b->cap_forces != forces)) {
            if (b->exec[gi]) {
                (void)hipGraphExecDestroy(b->exec[gi]);
                b->exec[gi] = nullptr;
            }
            if (b->graph[gi]) {
                (void)hipGraphDestroy(b->graph[gi]);
                b->graph[gi] = nullptr;
            }
            HYMD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = batch_enqueue(b, x, types, forces);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (st != HYMD_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HYMD_HIP(e);
            b->graph[gi] = g;
            HYMD_HIP(hipGraphInstantiate(&b->exec[gi], g, nullptr, nullptr, 0));
            b->cap_x[gi] = x;
            b->cap_types[gi] = types;
            if (forces) b->cap_forces = forces;
        }
        HYMD_HIP(hipGraphLaunch(b->exec[gi], s));
    } else {
        HYMD_TRY(batch_enqueue(b, x, types, forces));
    }
    HYMD_HIP(hipStreamSynchronize(s));
    const int np = T * (T + 1) / 2, nv = 2 * np + T;
    const double *hr = (const double *)b->h_result.ptr;
    for (int r = 0; r < R; ++r) std::memcpy(moments + (int64_t)r * nv, hr + (int64_t)b->slot_of[r] * b->row_len,
                                            sizeof(double) * nv);
    return HYMD_OK;
}

extern "C" int hymd_batch_moments(hymd_batch *b, const float *x, const int32_t *types, const double *sigma,
                                  double *moments, int32_t use_graph, void *stream) {
    return batch_run(b, x, types, sigma, nullptr, moments, nullptr, use_graph, stream);
}

extern "C" int hymd_batch_forces(hymd_batch *b, const float *x, const int32_t *types, const double *sigma,
                                 const double *coupling, double *moments, float *forces, int32_t use_graph,
                                 void *stream) {
    HYMD_REQUIRE(forces || (b && b->n == 0), "null forces");
    return batch_run(b, x, types, sigma, coupling, moments, forces, use_graph, stream);
}
