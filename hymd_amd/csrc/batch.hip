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

__device__ __forceinline__ int find_slot_by_row(const RepGeom *__restrict__ geo, int S, int row) {
    int lo = 0, hi = S;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (geo[mid].row_off <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlock) batch_key_kernel(const float *__restrict__ x, int64_t n,
                                                            const int64_t *__restrict__ poff, int R,
                                                            const int *__restrict__ slot_of,
                                                            const RepGeom *__restrict__ geo, int2 *__restrict__ keys,
                                                            int *__restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x; i0 < n;
         i0 += stride) {  // uniform per wavefront
        const int64_t i = i0 + threadIdx.x;
        int key = -1;
        if (i < n) {
            const int r = find_range(poff, R, i);
            const RepGeom &G = geo[slot_of[r]];
            const CicAxis cx = cic_axis((double)x[3 * i + 0], G.inv_h[0], G.mx);
            const CicAxis cy = cic_axis((double)x[3 * i + 1], G.inv_h[1], G.my);
            key = G.row_off + cx.i0 * G.my + cy.i0;
        }
        const int rank = rank_in_bin_all(key, counts);
        if (i < n) keys[i] = make_int2(key, rank);
    }
}

// One workgroup per global mesh row (all replicas): the row paint of the
// single-mesh path with the replica's geometry from the table.
__global__ void __launch_bounds__(kBlock) batch_paint_rows_kernel(const float4 *__restrict__ rec,
                                                                   const int *__restrict__ start,
                                                                   const RepGeom *__restrict__ geo, int S,
                                                                   PaintArgs pa, float *__restrict__ out) {
#pragma clang fp contract(off)
    extern __shared__ unsigned long long acc[];  // T * mz
    const int grow = (int)xcd_block(blockIdx.x, gridDim.x);
    const int slot = find_slot_by_row(geo, S, grow);
    const RepGeom &G = geo[slot];
    const int T = pa.n_types;
    const int row = grow - G.row_off;
    const int ix = row / G.my, iy = row - ix * G.my;
    for (int i = threadIdx.x; i < T * G.mz; i += blockDim.x) acc[i] = 0ull;
    int rs[4], rn[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = q >> 1, b = q & 1;
        const int bx = a ? (ix == 0 ? G.mx - 1 : ix - 1) : ix;
        const int by = b ? (iy == 0 ? G.my - 1 : iy - 1) : iy;
        const int br = G.row_off + bx * G.my + by;
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
        const int t = __float_as_int(r.w);
        if ((unsigned)t >= (unsigned)T) continue;
        row_deposit(r, q >> 1, q & 1, pa.mass[t], pa.scale, G.inv_h, G.mx, G.my, G.mz, acc + t * G.mz);
    }
    __syncthreads();
    float *o = out + G.grid_off + (int64_t)row * G.mz;
    for (int i = threadIdx.x; i < T * G.mz; i += blockDim.x) {
        const int t = i / G.mz, z = i - t * G.mz;
        o[(int64_t)t * G.g + z] = (float)((double)(long long)acc[i] * G.out_scale);
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
    DevBuf keys, perm, xs, counts, start, tiles;  // xs: binned records (x, y, z, type bits)
    DevBuf rho, spec, partials, result;
    HostBuf h_sigma2, h_result;
    // captured sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const float *cap_x = nullptr;
    const int32_t *cap_types = nullptr;
    ~hymd_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static int batch_enqueue(hymd_batch *b, const float *x, const int32_t *types) {
    hipStream_t s = b->stream;
    const int S = b->R;
    if (b->n > 0)
        batch_key_kernel<<<grid_for(b->n), kBlock, 0, s>>>(x, b->n, (const int64_t *)b->d_poff.ptr, b->R,
                                                           (const int *)b->d_slot_of.ptr,
                                                           (const RepGeom *)b->d_geo.ptr, (int2 *)b->keys.ptr,
                                                           (int *)b->counts.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan((int *)b->counts.ptr, (int *)b->start.ptr, b->rows, (int *)b->tiles.ptr, nullptr,
                            true, s));
    HYMD_TRY(launch_key_scatter(x, types, nullptr, b->n, (const int2 *)b->keys.ptr, (const int *)b->start.ptr,
                                (int *)b->perm.ptr, (float4 *)b->xs.ptr, nullptr, s));
    int max_mz = 0;
    for (const RepGeom &G : b->geo) max_mz = std::max(max_mz, G.mz);
    batch_paint_rows_kernel<<<b->rows, kBlock, sizeof(long long) * b->T * max_mz, s>>>(
        (const float4 *)b->xs.ptr, (const int *)b->start.ptr,
        (const RepGeom *)b->d_geo.ptr, S, b->pa, (float *)b->rho.ptr);
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
    const size_t nn = (size_t)std::max<int64_t>(b->n, 1);
    std::vector<int> nbs(R);
    std::vector<int64_t> row0(R);
    for (int sl = 0; sl < R; ++sl) {
        nbs[sl] = b->geo[sl].nb;
        row0[sl] = b->geo[sl].part_off;
    }
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
    hipStream_t s = b->stream;
    HYMD_HIP(hipMemcpyAsync(b->d_geo.ptr, b->geo.data(), sizeof(RepGeom) * R, hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_slot_of.ptr, b->slot_of.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_poff.ptr, b->poff.data(), sizeof(int64_t) * (R + 1), hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(b->d_nb.ptr, nbs.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
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

extern "C" int hymd_batch_moments(hymd_batch *b, const float *x, const int32_t *types, const double *sigma,
                                  double *moments, int32_t use_graph, void *stream) {
    HYMD_REQUIRE(b && sigma && moments && ((x && types) || b->n == 0), "null argument");
    const int R = b->R, T = b->T;
    double *hs = (double *)b->h_sigma2.ptr;
    for (int r = 0; r < R; ++r) {
        HYMD_REQUIRE(sigma[r] >= 0.0 && std::isfinite(sigma[r]), "replica %d: sigma must be finite and >= 0", r);
        hs[b->slot_of[r]] = sigma[r] * sigma[r];
    }
    hipStream_t s = b->stream;
    // inputs written on the caller's stream are ready before the batch reads them
    HYMD_HIP(hipEventRecord(b->ev_in, (hipStream_t)stream));
    HYMD_HIP(hipStreamWaitEvent(s, b->ev_in, 0));
    HYMD_HIP(hipMemcpyAsync(b->d_sigma2.ptr, hs, sizeof(double) * R, hipMemcpyHostToDevice, s));
    if (use_graph) {
        if (!b->exec || b->cap_x != x || b->cap_types != types) {
            if (b->exec) {
                (void)hipGraphExecDestroy(b->exec);
                b->exec = nullptr;
            }
            if (b->graph) {
                (void)hipGraphDestroy(b->graph);
                b->graph = nullptr;
            }
            HYMD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const int st = batch_enqueue(b, x, types);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (st != HYMD_OK) {
                if (g) (void)hipGraphDestroy(g);
                return st;
            }
            HYMD_HIP(e);
            b->graph = g;
            HYMD_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
            b->cap_x = x;
            b->cap_types = types;
        }
        HYMD_HIP(hipGraphLaunch(b->exec, s));
    } else {
        HYMD_TRY(batch_enqueue(b, x, types));
    }
    HYMD_HIP(hipStreamSynchronize(s));
    const int np = T * (T + 1) / 2, nv = 2 * np + T;
    const double *hr = (const double *)b->h_result.ptr;
    for (int r = 0; r < R; ++r) std::memcpy(moments + (int64_t)r * nv, hr + (int64_t)b->slot_of[r] * b->row_len,
                                            sizeof(double) * nv);
    return HYMD_OK;
}
