// CSVR thermostat (thermostat.py:111-219) and centre-of-mass momentum removal
// (thermostat.py:12-15) on device FP64 velocities, without a host round trip.
//
// The reference walks the coupling groups in order; each group either
//  * (COM removal, more than one particle) rescales the COM-free part of its
//    own velocities from the kinetic energy of that part, or
//  * (otherwise) rescales ALL velocities from the kinetic energy of ALL
//    particles.
// Every such update is affine in the velocities it starts from, so the whole
// call is one affine map per particle class v -> a_c v + b_c, where class c is
// the COM group that holds the particle's type (or "none").  Three launches:
//   moments (grid)  per class sum(v - s_c) and sum|v - s_c|^2, over all
//                   particles sum|v|^2, per workgroup (column-major partials)
//   solve   (1 WG)  fixed-order sums of the partials; the reference's group
//                   loop on one lane -> a_c, b_c, the work, the new s_c
//   scale   (grid)  v <- a_c v + b_c
// The shift s_c is the class's COM velocity from the previous call (zero at
// first), so the COM-free sum of squares sum|v - s|^2 - |sum(v - s)|^2 / n
// loses no digits to a large COM velocity.  The random numbers (R, chi^2
// with 3 N_g - 1 degrees of freedom) come from the caller, drawn in the
// reference's order; nothing is read back.
#include <cmath>

#include "common.h"

namespace hymd {

constexpr int kThermoMaxGroups = HYMD_MAX_TYPES;       // one group per type at most
constexpr int kThermoClasses = kThermoMaxGroups + 1;   // + "none"
constexpr int kThermoMaxBlocks = 1024;
constexpr int kThermoMaxNv = 4 * kThermoMaxGroups + 1;

// workspace layout (doubles)
constexpr int kWsCoef = 0;                                  // [class][a, bx, by, bz]
constexpr int kWsShift = kWsCoef + 4 * kThermoClasses;      // [class][3] COM of the last call
constexpr int kWsSums = kWsShift + 3 * kThermoMaxGroups;    // [nv] reduced sums
constexpr int kWsPart = kWsSums + kThermoMaxNv;             // [nv][blocks]
constexpr int64_t kWsDoubles = kWsPart + (int64_t)kThermoMaxBlocks * kThermoMaxNv;

// class of a type: 4 bits per type packed in a 64-bit word (16 types)
__device__ __forceinline__ int type_class(uint64_t tab, int t) { return (int)((tab >> (4 * t)) & 15u); }

// G = COM classes rounded up to a power of two (0: only sum |v|^2).  Columns:
// 4 j + (0,1,2) = sum(v - s_j), 4 j + 3 = sum|v - s_j|^2, 4 G = sum|v|^2.
// types == NULL: every particle in class 0; shift == NULL: s = 0.
template <int G>
__global__ void __launch_bounds__(kBlock) csvr_moments_kernel(const double *__restrict__ v,
                                                               const int32_t *__restrict__ types, int64_t n,
                                                               uint64_t cls_tab, const double *__restrict__ shift,
                                                               const float *__restrict__ fk, double ck,
                                                               double *__restrict__ partials) {
    constexpr int NV = 4 * G + 1;
    double acc[NV], sh[3 * G > 0 ? 3 * G : 1];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
#pragma unroll
    for (int j = 0; j < 3 * G; ++j) sh[j] = shift ? shift[j] : 0.0;
    // four particles per lane per round, their loads issued together
    constexpr int U = 4;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += U * stride) {
        double px[U], py[U], pz[U];
        int pc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = i0 + u * stride;
            const bool ok = i < n;
            px[u] = ok ? v[3 * i] : 0.0;
            py[u] = ok ? v[3 * i + 1] : 0.0;
            pz[u] = ok ? v[3 * i + 2] : 0.0;
            if (fk && ok) {  // the pending outer half kick, as the kick kernel adds it
                px[u] = px[u] + ck * (double)fk[3 * i];
                py[u] = py[u] + ck * (double)fk[3 * i + 1];
                pz[u] = pz[u] + ck * (double)fk[3 * i + 2];
            }
            pc[u] = !ok ? -1 : (G > 0 && types ? type_class(cls_tab, types[i]) : 0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double vx = px[u], vy = py[u], vz = pz[u];
            acc[4 * G] += vx * vx + vy * vy + vz * vz;
            if constexpr (G > 0) {
                const int c = pc[u];
                double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
                for (int j = 0; j < G; ++j)
                    if (c == j) sx = sh[3 * j], sy = sh[3 * j + 1], sz = sh[3 * j + 2];
                const double dx = vx - sx, dy = vy - sy, dz = vz - sz;
                const double d2 = dx * dx + dy * dy + dz * dz;
#pragma unroll
                for (int j = 0; j < G; ++j) {
                    const bool sel = c == j;
                    acc[4 * j] += sel ? dx : 0.0;
                    acc[4 * j + 1] += sel ? dy : 0.0;
                    acc[4 * j + 2] += sel ? dz : 0.0;
                    acc[4 * j + 3] += sel ? d2 : 0.0;
                }
            }
        }
    }
    // per-workgroup sums -> column j, row blockIdx.x
    constexpr int kWaves = kBlock / 64;
    __shared__ double red[NV][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double t = acc[j];
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
        if (lane == 0) red[j][wid] = t;
    }
    __syncthreads();
    for (int j = threadIdx.x; j < NV; j += blockDim.x) {
        double t = 0.0;
        for (int w = 0; w < kWaves; ++w) t += red[j][w];
        partials[(int64_t)j * gridDim.x + blockIdx.x] = t;
    }
}


// Fixed-order column sums of the partials (one wavefront per column, lanes
// over rows, then a shuffle tree) into sums[] (LDS) and ws[kWsSums..].
__device__ void sum_columns(const double *__restrict__ part, int nb, int nv, double *sums, double *__restrict__ out) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int j = wid; j < nv; j += nw) {
        double t = lane_strided_sum(part + (int64_t)j * nb, nb, 1, lane);
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
        if (lane == 0) {
            sums[j] = t;
            out[j] = t;
        }
    }
    __syncthreads();
}

// One workgroup: the sums, then the reference's per-group loop on one lane,
// tracking the affine map of every class and the current sum|v|^2 of all
// particles.
__global__ void __launch_bounds__(1024) csvr_solve_kernel(CsvrGroups gp, double *__restrict__ ws,
                                                           double *__restrict__ work) {
    __shared__ double sums[kThermoMaxNv];
    __shared__ double com[3 * kThermoMaxGroups], s2[kThermoMaxGroups];  // lane 0's tables
    const int nv = 4 * gp.g2 + 1;
    sum_columns(ws + kWsPart, gp.nb, nv, sums, ws + kWsSums);
    if (threadIdx.x != 0) return;
    double *coef = ws + kWsCoef;
    double *shift = ws + kWsShift;
    const int nc = gp.n_com;
    for (int g = 0; g < gp.n_groups; ++g) {
        const int k = gp.cls[g];
        if (k < 0) continue;
        const double ng = gp.n[g];
        const double mx = sums[4 * k] / ng, my = sums[4 * k + 1] / ng, mz = sums[4 * k + 2] / ng;
        // sum |v - com|^2 = sum |v - s|^2 - n |com - s|^2
        s2[k] = sums[4 * k + 3] - ng * (mx * mx + my * my + mz * mz);
        com[3 * k] = shift[3 * k] + mx;
        com[3 * k + 1] = shift[3 * k + 1] + my;
        com[3 * k + 2] = shift[3 * k + 2] + mz;
    }
    for (int c = 0; c <= nc; ++c) {
        coef[4 * c] = 1.0;
        coef[4 * c + 1] = coef[4 * c + 2] = coef[4 * c + 3] = 0.0;
    }
    double total = sums[4 * gp.g2];  // sum |v|^2 over all particles, current
    double w = *work;
    const double c = gp.c;
    for (int g = 0; g < gp.n_groups; ++g) {
        const int k = gp.cls[g];
        const double ng = gp.n[g];
        double K;
        if (k >= 0) {
            const double a = coef[4 * k];
            K = gp.half_mass * (a * a * s2[k]);
        } else {
            K = gp.half_mass * total;
        }
        const double Kt = gp.kt_per_particle * ng;
        const double nf = 3.0 * ng;
        const double R = gp.R[g];
        const double alpha2 = c + (1.0 - c) * (gp.S[g] + R * R) * Kt / (nf * K) +
                              2.0 * R * sqrt(c * (1.0 - c) * Kt / (nf * K));
        w += K * (alpha2 - 1.0);
        const double alpha = sqrt(alpha2);
        if (k >= 0) {
            // group velocities a v + b; COM-free part a (v - com): scale it by alpha
            const double a = coef[4 * k];
            const double f = a * (1.0 - alpha);
            for (int d = 0; d < 3; ++d) coef[4 * k + 1 + d] += f * com[3 * k + d];
            coef[4 * k] = alpha * a;
            total += (alpha2 - 1.0) * a * a * s2[k];
        } else {
            for (int q = 0; q < 4 * (nc + 1); ++q) coef[q] *= alpha;
            total *= alpha2;
        }
    }
    // next call's shifts: the classes' COM velocities after this call
    // (any finite shift is exact; a non-finite one would poison later calls)
    for (int k = 0; k < nc; ++k)
        for (int d = 0; d < 3; ++d) {
            const double s_new = coef[4 * k] * com[3 * k + d] + coef[4 * k + 1 + d];
            shift[3 * k + d] = isfinite(s_new) ? s_new : 0.0;
        }
    *work = w;
}

__global__ void __launch_bounds__(kBlock) csvr_scale_kernel(double *__restrict__ v,
                                                             const int32_t *__restrict__ types, int64_t n,
                                                             uint64_t cls_tab, int n_com,
                                                             const float *__restrict__ fk, double ck,
                                                             const double *__restrict__ ws) {
    __shared__ double coef[4 * kThermoClasses];
    for (int q = threadIdx.x; q < 4 * (n_com + 1); q += blockDim.x) coef[q] = ws[kWsCoef + q];
    __syncthreads();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int c = n_com > 0 ? type_class(cls_tab, types[i]) : 0;
        const double a = coef[4 * c];
        double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
        if (fk) {
            vx = vx + ck * (double)fk[3 * i];
            vy = vy + ck * (double)fk[3 * i + 1];
            vz = vz + ck * (double)fk[3 * i + 2];
        }
        v[3 * i] = a * vx + coef[4 * c + 1];
        v[3 * i + 1] = a * vy + coef[4 * c + 2];
        v[3 * i + 2] = a * vz + coef[4 * c + 3];
    }
}

// sum(v) of every particle: the moments kernel's first three columns (G = 1,
// no shift), summed by one workgroup
__global__ void __launch_bounds__(1024) com_sum_kernel(double *__restrict__ ws, int nb) {
    __shared__ double sums[5];
    sum_columns(ws + kWsPart, nb, 5, sums, ws + kWsSums);
}

// v -= sum(v) / n_particles (the quotient first, as thermostat.py:14)
__global__ void __launch_bounds__(kBlock) com_subtract_kernel(double *__restrict__ v, int64_t n,
                                                               const double *__restrict__ sum1, double n_particles) {
    const double q0 = sum1[0] / n_particles, q1 = sum1[1] / n_particles, q2 = sum1[2] / n_particles;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        v[3 * i] -= q0;
        v[3 * i + 1] -= q1;
        v[3 * i + 2] -= q2;
    }
}

static int launch_moments(int g, int nb, const double *v, const int32_t *types, int64_t n, uint64_t tab,
                          const double *shift, const float *fk, double ck, double *part, hipStream_t s) {
    switch (g) {
        case 0: csvr_moments_kernel<0><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        case 1: csvr_moments_kernel<1><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        case 2: csvr_moments_kernel<2><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        case 4: csvr_moments_kernel<4><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        case 8: csvr_moments_kernel<8><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        case 16: csvr_moments_kernel<16><<<nb, kBlock, 0, s>>>(v, types, n, tab, shift, fk, ck, part); break;
        default: set_error("internal: bad class count %d", g); return HYMD_ERR_INVALID;
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

static int pow2_classes(int n) {
    int g = 0;
    while (g < n) g = g ? 2 * g : 1;
    return g;
}

}  // namespace hymd

using namespace hymd;

extern "C" int64_t hymd_thermostat_workspace_bytes(void) { return kWsDoubles * (int64_t)sizeof(double); }

namespace hymd {

int csvr_setup(CsvrSetup *cs, int32_t n_types, const int32_t *type_group, int32_t n_groups, const int64_t *group_n,
               double mass, double gas_constant, double target_temperature, double c, int32_t remove_com) {
    HYMD_REQUIRE(cs, "NULL setup");
    HYMD_REQUIRE(n_types >= 1 && n_types <= HYMD_MAX_TYPES, "n_types must be in [1, %d]", HYMD_MAX_TYPES);
    HYMD_REQUIRE(n_groups >= 1 && n_groups <= kThermoMaxGroups, "n_groups must be in [1, %d]", kThermoMaxGroups);
    HYMD_REQUIRE(type_group && group_n, "NULL argument");
    HYMD_REQUIRE(mass > 0.0 && c >= 0.0 && c <= 1.0, "need mass > 0 and c in [0, 1]");
    for (int g = 0; g < n_groups; ++g)
        HYMD_REQUIRE(group_n[g] > 0, "thermostat coupling group %d has no particles", g);
    // COM classes: groups with COM removal and more than one particle
    CsvrGroups &gp = cs->gp;
    gp = CsvrGroups{};
    gp.n_groups = n_groups;
    gp.half_mass = 0.5 * mass;
    gp.kt_per_particle = 1.5 * gas_constant * target_temperature;
    gp.c = c;
    int n_com = 0;
    for (int g = 0; g < n_groups; ++g) {
        gp.cls[g] = (remove_com && group_n[g] > 1) ? n_com++ : -1;
        gp.n[g] = (double)group_n[g];
    }
    gp.n_com = n_com;
    uint64_t tab = 0;
    int typed_in_com = 0;
    for (int t = 0; t < n_types; ++t) {
        const int g = type_group[t];
        HYMD_REQUIRE(g >= -1 && g < n_groups, "type_group[%d] = %d is out of range", t, g);
        const int cls = (g >= 0 && gp.cls[g] >= 0) ? gp.cls[g] : n_com;
        typed_in_com += cls < n_com;
        tab |= (uint64_t)(cls & 15) << (4 * t);
    }
    // a class index must fit 4 bits: with 16 COM classes every type is in one of them
    HYMD_REQUIRE(n_com < 16 || typed_in_com == n_types, "internal: class table overflow");
    cs->tab = tab;
    gp.g2 = pow2_classes(n_com);
    return HYMD_OK;
}

int csvr_enqueue(const CsvrSetup &cs, double *v, const int32_t *types, int64_t n, const double *gaussians,
                 const double *chi_squares, const float *fkick, double ckick, double *work, double *ws,
                 hipStream_t s) {
    CsvrGroups gp = cs.gp;
    for (int g = 0; g < gp.n_groups; ++g) {
        gp.R[g] = gaussians[g];
        gp.S[g] = chi_squares[g];
    }
    const int nb = grid_for(n > 0 ? n : 1, kBlock, kThermoMaxBlocks);
    gp.nb = nb;
    HYMD_TRY(launch_moments(gp.g2, nb, v, types, n, cs.tab, ws + kWsShift, fkick, ckick, ws + kWsPart, s));
    csvr_solve_kernel<<<1, 1024, 0, s>>>(gp, ws, work);
    HYMD_HIP(hipGetLastError());
    if (n > 0) csvr_scale_kernel<<<nb, kBlock, 0, s>>>(v, types, n, cs.tab, gp.n_com, fkick, ckick, ws);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int com_enqueue(double *v, int64_t n, int64_t n_particles, double *ws, hipStream_t s) {
    HYMD_REQUIRE(n >= 0 && (n == 0 || v) && n_particles > 0, "bad argument");
    if (n == 0) return HYMD_OK;
    const int nb = grid_for(n, kBlock, kThermoMaxBlocks);
    // one class holding every particle, no shift: columns 0-2 = sum(v)
    HYMD_TRY(launch_moments(1, nb, v, nullptr, n, 0, nullptr, nullptr, 0.0, ws + kWsPart, s));
    com_sum_kernel<<<1, 1024, 0, s>>>(ws, nb);
    HYMD_HIP(hipGetLastError());
    com_subtract_kernel<<<nb, kBlock, 0, s>>>(v, n, ws + kWsSums, (double)n_particles);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int64_t thermostat_workspace_doubles() { return kWsDoubles; }

}  // namespace hymd

extern "C" int hymd_csvr_thermostat(double *v, const int32_t *types, int64_t n, int32_t n_types,
                                    const int32_t *type_group, int32_t n_groups, const int64_t *group_n,
                                    const double *gaussians, const double *chi_squares, double mass,
                                    double gas_constant, double target_temperature, double c, int32_t remove_com,
                                    double *work, void *workspace, int64_t workspace_bytes, void *stream) {
    HYMD_REQUIRE(n >= 0 && (n == 0 || (v && types)), "bad velocity/type arrays");
    HYMD_REQUIRE(gaussians && chi_squares && work, "NULL argument");
    HYMD_REQUIRE(workspace && workspace_bytes >= kWsDoubles * (int64_t)sizeof(double),
                 "workspace must hold hymd_thermostat_workspace_bytes() bytes");
    CsvrSetup cs;
    HYMD_TRY(csvr_setup(&cs, n_types, type_group, n_groups, group_n, mass, gas_constant, target_temperature, c,
                        remove_com));
    return csvr_enqueue(cs, v, types, n, gaussians, chi_squares, nullptr, 0.0, work, (double *)workspace,
                        (hipStream_t)stream);
}

extern "C" int hymd_cancel_com_momentum(double *v, int64_t n, int64_t n_particles, void *workspace,
                                        int64_t workspace_bytes, void *stream) {
    HYMD_REQUIRE(workspace && workspace_bytes >= kWsDoubles * (int64_t)sizeof(double),
                 "workspace must hold hymd_thermostat_workspace_bytes() bytes");
    return com_enqueue(v, n, n_particles, (double *)workspace, (hipStream_t)stream);
}
