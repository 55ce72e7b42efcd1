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

#include "fft.h"

namespace hymd {

// ---------------------------------------------------------------------------
// paint
// ---------------------------------------------------------------------------
struct PaintArgs {
    int n_types;
    double scale;  // 2^HYMD_FIXED_SHIFT
    double mass[HYMD_MAX_TYPES];
};

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
// Spatial sort: particles binned by their CIC base cell (counting sort), then
// paint as a gather over cells and readout in cell order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int base_cell(const float *__restrict__ x, int64_t i, const Mesh &m) {
    const CicAxis cx = cic_axis((double)x[3 * i + 0], m.inv_h[0], m.mx);
    const CicAxis cy = cic_axis((double)x[3 * i + 1], m.inv_h[1], m.my);
    const CicAxis cz = cic_axis((double)x[3 * i + 2], m.inv_h[2], m.mz);
    return (cx.i0 * m.my + cy.i0) * m.mz + cz.i0;
}

__global__ void __launch_bounds__(kBlock) cell_key_kernel(const float *__restrict__ x, int64_t n, Mesh m,
                                                           int *__restrict__ keys, int *__restrict__ counts) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int key = base_cell(x, i, m);
        keys[i] = key;
        atomicAdd(counts + key, 1);
    }
}

__global__ void __launch_bounds__(kBlock) cell_scatter_kernel(const float *__restrict__ x,
                                                               const int32_t *__restrict__ types,
                                                               const float *__restrict__ w, int64_t n,
                                                               const int *__restrict__ keys,
                                                               int *__restrict__ cursor, int *__restrict__ perm,
                                                               float *__restrict__ xs, int32_t *__restrict__ ts,
                                                               float *__restrict__ ws) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int pos = atomicAdd(cursor + keys[i], 1);
        perm[pos] = (int)i;
        xs[3 * (int64_t)pos] = x[3 * i];
        xs[3 * (int64_t)pos + 1] = x[3 * i + 1];
        xs[3 * (int64_t)pos + 2] = x[3 * i + 2];
        ts[pos] = types ? types[i] : 0;
        if (ws) ws[pos] = w[i];
    }
}

// One thread per cell: sum the fixed-point corner deposits of the particles
// binned in the 8 cells whose upper corners include this one (the same int64
// values paint_fixed_kernel adds atomically, so the sums are identical and
// independent of order), then write rho = sum * 2^-32 / dV per type.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) paint_gather_kernel(const float *__restrict__ xs,
                                                               const int32_t *__restrict__ ts,
                                                               const float *__restrict__ ws,
                                                               const int *__restrict__ start, Mesh m, PaintArgs pa,
                                                               double out_scale, float *__restrict__ out,
                                                               long long *__restrict__ out_fixed) {
#pragma clang fp contract(off)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t cell = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; cell < m.g; cell += stride) {
        const int iz = (int)(cell % m.mz);
        const int iy = (int)((cell / m.mz) % m.my);
        const int ix = (int)(cell / ((int64_t)m.mz * m.my));
        long long acc[MAXT];
#pragma unroll
        for (int t = 0; t < MAXT; ++t) acc[t] = 0;
        for (int c = 0; c < 8; ++c) {
            const int a = c >> 2, b = (c >> 1) & 1, d = c & 1;
            const int bx = a ? (ix == 0 ? m.mx - 1 : ix - 1) : ix;
            const int by = b ? (iy == 0 ? m.my - 1 : iy - 1) : iy;
            const int bz = d ? (iz == 0 ? m.mz - 1 : iz - 1) : iz;
            const int bin = (bx * m.my + by) * m.mz + bz;
            for (int q = start[bin]; q < start[bin + 1]; ++q) {
                const int t = ts[q];
                if ((unsigned)t >= (unsigned)pa.n_types) continue;
                const double mass = ws ? (double)ws[q] : pa.mass[t];
                const CicAxis cx = cic_axis((double)xs[3 * (int64_t)q + 0], m.inv_h[0], m.mx);
                const CicAxis cy = cic_axis((double)xs[3 * (int64_t)q + 1], m.inv_h[1], m.my);
                const CicAxis cz = cic_axis((double)xs[3 * (int64_t)q + 2], m.inv_h[2], m.mz);
                long long qc;
                if (c < 7) {
                    const double wgt = ((a ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0)) * (d ? cz.w1 : cz.w0);
                    qc = __double2ll_rn((wgt * mass) * pa.scale);
                } else {
                    long long s7 = 0;
#pragma unroll
                    for (int e = 0; e < 7; ++e) {
                        const int ea = e >> 2, eb = (e >> 1) & 1, ed = e & 1;
                        const double wgt = ((ea ? cx.w1 : cx.w0) * (eb ? cy.w1 : cy.w0)) * (ed ? cz.w1 : cz.w0);
                        s7 += __double2ll_rn((wgt * mass) * pa.scale);
                    }
                    qc = __double2ll_rn(mass * pa.scale) - s7;
                }
#pragma unroll
                for (int u = 0; u < MAXT; ++u)
                    if (u == t) acc[u] += qc;
            }
        }
#pragma unroll
        for (int t = 0; t < MAXT; ++t) {
            if (t >= pa.n_types) break;
            out[(int64_t)t * m.g + cell] = (float)((double)acc[t] * out_scale);
            if (out_fixed) out_fixed[(int64_t)t * m.g + cell] = acc[t];
        }
    }
}

// ---------------------------------------------------------------------------
// readout: out[i, c] = w_i * sum_corners W * grid[type_i * ncomp + c][corner]
// ---------------------------------------------------------------------------
// perm != NULL: particle i of the (sorted) inputs is written to row perm[i]
template <typename PosT>
__global__ void __launch_bounds__(kBlock) readout_kernel(const float *__restrict__ grids, int ncomp,
                                                          const PosT *__restrict__ x,
                                                          const int32_t *__restrict__ types,
                                                          const float *__restrict__ scale_per_particle,
                                                          int64_t n, int n_types, Mesh m,
                                                          float *__restrict__ out, const int *__restrict__ perm) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const int t = types ? types[i] : 0;
        const int64_t o = perm ? (int64_t)perm[i] : i;
        if ((unsigned)t >= (unsigned)n_types) {
            for (int c = 0; c < ncomp; ++c) out[ncomp * o + c] = 0.f;
            continue;
        }
        const CicAxis cx = cic_axis((double)x[3 * i + 0], m.inv_h[0], m.mx);
        const CicAxis cy = cic_axis((double)x[3 * i + 1], m.inv_h[1], m.my);
        const CicAxis cz = cic_axis((double)x[3 * i + 2], m.inv_h[2], m.mz);
        double acc[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int a = c >> 2, b = (c >> 1) & 1, d = c & 1;
            const double w = ((a ? cx.w1 : cx.w0) * (b ? cy.w1 : cy.w0)) * (d ? cz.w1 : cz.w0);
            const int64_t idx =
                ((int64_t)(a ? cx.i1 : cx.i0) * m.my + (b ? cy.i1 : cy.i0)) * m.mz + (d ? cz.i1 : cz.i0);
            for (int q = 0; q < ncomp; ++q) acc[q] += w * (double)grids[((int64_t)t * ncomp + q) * m.g + idx];
        }
        const double s = scale_per_particle ? (double)scale_per_particle[i] : 1.0;
        for (int q = 0; q < ncomp; ++q) out[ncomp * o + q] = (float)(acc[q] * s);
    }
}

// ---------------------------------------------------------------------------
// k-space helpers
// ---------------------------------------------------------------------------
struct KGeom {
    int mx, my, mz, zh;
    int64_t gc;
    double kfac[3];  // 2 pi / L_d
};

__device__ __forceinline__ void kvec(const KGeom &g, int64_t kk, double k[3], double kd[3], double &wgt) {
    const int64_t plane = (int64_t)g.my * g.zh;
    const int ix = (int)(kk / plane);
    const int64_t r = kk - (int64_t)ix * plane;
    const int iy = (int)(r / g.zh);
    const int iz = (int)(r - (int64_t)iy * g.zh);
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

// Spectral moments of the filtered densities phi~_t = H rho_t / G (pmesh
// normalisation), summed over the full spectrum through the half spectrum
// weights: P_ts = sum_k w Re(phi~_t^* phi~_s), Q_ts = sum_k w |k|^2 Re(...),
// t <= s packed row-major, then phi~_t(0).  Every quadratic Hamiltonian's
// energy and its derivatives in chi, kappa and sigma follow from them.
template <int MAXT>
__global__ void __launch_bounds__(kBlock) field_moments_kernel(const float2 *__restrict__ spec, KGeom g, int T,
                                                                double sigma2, double inv_g,
                                                                double *__restrict__ partials) {
    constexpr int NP = MAXT * (MAXT + 1) / 2;
    double acc[2 * NP + MAXT];
#pragma unroll
    for (int j = 0; j < 2 * NP + MAXT; ++j) acc[j] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; kk < g.gc; kk += stride) {
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
    // one row of 2 NP + MAXT doubles per block (the first 2 T(T+1)/2 + T are used)
    block_reduce_store<2 * NP + MAXT>(out, partials + (int64_t)blockIdx.x * (2 * NP + MAXT));
}

struct PmeK {
    double sigma2;
    double inv_g;
    double four_pi_elec;  // 4 pi k_e / eps_r
};

// PME (field.py:363-397): rho~ = H rho/G; psi = 4 pi elec rho~/|k|^2 (|k|^2 := 1
// at k = 0); E_d = -i k_d psi.  Energy partials: 1/2 wgt Re(rho^* psi).
__global__ void __launch_bounds__(kBlock) pme_kspace_kernel(const float2 *__restrict__ spec,
                                                             float2 *__restrict__ out, int write_psi, KGeom g,
                                                             PmeK p, double *__restrict__ partials) {
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
    DevBuf paint;     // max_types x G int64 (zero between calls)
    DevBuf spec;      // max_types x Gc complex
    DevBuf grad;      // max(3 max_types, 4) x Gc complex
    DevBuf cscratch;  // max_types x Gc complex (phi / v_ext inverse transforms)
    DevBuf rscratch;  // 3 x G float (PME field when the caller passes none)
    DevBuf partials;  // energy partial sums
    DevBuf result;    // reduced doubles
    HostBuf host;     // pinned readback
    int kblocks;      // blocks of the k-space passes
    // spatial sort of the last painted particle set (cell-binned order)
    DevBuf keys, perm, xs, ts, ws;  // per particle
    DevBuf counts, start, cursor, tiles;  // per cell (counts kept zero between sorts)
    int64_t sorted_n = 0;
    bool sorted_weights = false;
};

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
    if (st == HYMD_OK) st = pm->partials.reserve(sizeof(double) * (pm->kblocks + 16));
    if (st == HYMD_OK) st = pm->result.reserve(sizeof(double) * 16);
    if (st == HYMD_OK) st = pm->host.reserve(sizeof(double) * 16);
    if (st == HYMD_OK) st = pm->counts.reserve(sizeof(int) * (m.g + 2));
    if (st == HYMD_OK) st = pm->start.reserve(sizeof(int) * (m.g + 1));
    if (st == HYMD_OK) st = pm->cursor.reserve(sizeof(int) * m.g);
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

extern "C" int hymd_pm_set_box(hymd_pm *pm, const double box[3]) {
    HYMD_REQUIRE(pm && box, "null argument");
    int32_t mesh[3] = {pm->mesh.mx, pm->mesh.my, pm->mesh.mz};
    return set_mesh(pm, mesh, box);
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

// Bin particles by CIC base cell: perm[pos] = original index, xs/ts/ws the
// sorted positions, types (0 when types == NULL) and weights (if w != NULL),
// start[c] .. start[c+1] the particles based in cell c.
static int sort_particles(hymd_pm *pm, const float *x, const int32_t *types, const float *w, int64_t n) {
    const Mesh &m = pm->mesh;
    hipStream_t s = pm->stream;
    HYMD_REQUIRE(n < (int64_t(1) << 31) && m.g < (int64_t(1) << 31), "too many particles or cells to sort");
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    HYMD_TRY(pm->keys.reserve(sizeof(int) * nn));
    HYMD_TRY(pm->perm.reserve(sizeof(int) * nn));
    HYMD_TRY(pm->xs.reserve(sizeof(float) * 3 * nn));
    HYMD_TRY(pm->ts.reserve(sizeof(int32_t) * nn));
    if (w) HYMD_TRY(pm->ws.reserve(sizeof(float) * nn));
    if (n > 0)
        cell_key_kernel<<<grid_for(n), kBlock, 0, s>>>(x, n, m, (int *)pm->keys.ptr, (int *)pm->counts.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan((int *)pm->counts.ptr, (int *)pm->start.ptr, m.g, (int *)pm->tiles.ptr,
                            (int *)pm->cursor.ptr, true, s));
    if (n > 0)
        cell_scatter_kernel<<<grid_for(n), kBlock, 0, s>>>(x, types, w, n, (const int *)pm->keys.ptr,
                                                           (int *)pm->cursor.ptr, (int *)pm->perm.ptr,
                                                           (float *)pm->xs.ptr, (int32_t *)pm->ts.ptr,
                                                           w ? (float *)pm->ws.ptr : nullptr);
    HYMD_HIP(hipGetLastError());
    pm->sorted_n = n;
    pm->sorted_weights = (w != nullptr);
    return HYMD_OK;
}

// Gather paint of the sorted particles into float grids rho_t = paint/dV
// (T grids), optionally also the int64 fixed-point sums.
static int paint_sorted(hymd_pm *pm, int T, const double *mass, float *out, long long *out_fixed) {
    PaintArgs pa;
    pa.n_types = T;
    pa.scale = std::ldexp(1.0, HYMD_FIXED_SHIFT);
    for (int t = 0; t < HYMD_MAX_TYPES; ++t) pa.mass[t] = (mass && t < T) ? mass[t] : 1.0;
    const double out_scale = std::ldexp(1.0, -HYMD_FIXED_SHIFT) / cell_volume(pm);
    const float *ws = pm->sorted_weights ? (const float *)pm->ws.ptr : nullptr;
    const Mesh &m = pm->mesh;
    const int nb = grid_for(m.g);
    const float *xs = (const float *)pm->xs.ptr;
    const int32_t *ts = (const int32_t *)pm->ts.ptr;
    const int *st = (const int *)pm->start.ptr;
    if (T <= 4)
        paint_gather_kernel<4><<<nb, kBlock, 0, pm->stream>>>(xs, ts, ws, st, m, pa, out_scale, out, out_fixed);
    else if (T <= 8)
        paint_gather_kernel<8><<<nb, kBlock, 0, pm->stream>>>(xs, ts, ws, st, m, pa, out_scale, out, out_fixed);
    else
        paint_gather_kernel<16><<<nb, kBlock, 0, pm->stream>>>(xs, ts, ws, st, m, pa, out_scale, out, out_fixed);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// Readout of ncomp grids per type at the last sorted particle set, written to
// the particles' original rows (scaled by their weights if the sort had them).
static int readout_sorted(hymd_pm *pm, int ncomp, int n_types, const float *grids, float *out, bool weighted) {
    const int64_t n = pm->sorted_n;
    if (n > 0)
        readout_kernel<float><<<grid_for(n), kBlock, 0, pm->stream>>>(
            grids, ncomp, (const float *)pm->xs.ptr, (const int32_t *)pm->ts.ptr,
            weighted ? (const float *)pm->ws.ptr : nullptr, n, n_types, pm->mesh, out, (const int *)pm->perm.ptr);
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
    HYMD_TRY(fft_inverse_xy(pm->fft, (float2 *)pm->cscratch.ptr, batch, pm->stream));
    return fft_c2r_z(pm->fft, (const float2 *)pm->cscratch.ptr, out, batch, 1.f, pm->stream);
}

extern "C" int hymd_readout(hymd_pm *pm, const float *grid, const float *x, int64_t n, float *out) {
    HYMD_REQUIRE(pm && grid && out && (x || n == 0) && n >= 0, "bad argument");
    if (n > 0)
        readout_kernel<float><<<grid_for(n), kBlock, 0, pm->stream>>>(grid, 1, x, nullptr, nullptr, n, 1, pm->mesh,
                                                                        out, nullptr);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

extern "C" int hymd_update_field(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                                 const double *coupling, const double *shift, double constant,
                                 const double *mass, double sigma, float *force_mesh, float *phi,
                                 void *phi_fourier, float *v_ext, double *energy) {
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
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
    HYMD_TRY(paint_sorted(pm, T, mass, rho, nullptr));
    // 2. r2c
    HYMD_TRY(fft_r2c_real(pm->fft, rho, spec, T, s));
    // 3. fused filter / W' / filter / gradient
    FieldK p = field_params(T, coupling, shift, constant, sigma, m);
    p.want_energy = energy ? 1 : 0;
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
    HYMD_TRY(fft_inverse_xy(pm->fft, grad, 3 * T, s));
    HYMD_TRY(fft_c2r_z(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
    // 5. optional real-space outputs
    if (v_ext) {
        HYMD_TRY(fft_inverse_xy(pm->fft, spec, T, s));
        HYMD_TRY(fft_c2r_z(pm->fft, spec, v_ext, T, 1.f, s));
    }
    if (phi) {
        float2 *work = phif;
        if (phi_fourier) {  // keep the caller's spectra intact
            HYMD_TRY(pm->cscratch.reserve(sizeof(float2) * m.gc * T));
            HYMD_HIP(hipMemcpyAsync(pm->cscratch.ptr, phi_fourier, sizeof(float2) * m.gc * T,
                                    hipMemcpyDeviceToDevice, s));
            work = (float2 *)pm->cscratch.ptr;
        }
        HYMD_TRY(fft_inverse_xy(pm->fft, work, T, s));
        HYMD_TRY(fft_c2r_z(pm->fft, work, phi, T, 1.f, s));
    }
    if (energy) {
        double e = 0.0;
        HYMD_TRY(read_result(pm, kb, 1, &e));
        *energy = e * (m.box[0] * m.box[1] * m.box[2]);
    }
    return HYMD_OK;
}

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
    const int row = (T <= 4) ? 2 * 10 + 4 : 2 * 36 + 8;  // 2 NP + MAXT of the instance
    const int nb = std::min(pm->kblocks, 1024);
    DevBuf part;
    HYMD_TRY(part.reserve(sizeof(double) * (size_t)row * nb));
    HYMD_TRY(pm->result.reserve(sizeof(double) * row));
    HYMD_TRY(pm->host.reserve(sizeof(double) * row));
    if (T <= 4)
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
        HYMD_TRY(fft_inverse_xy(pm->fft, grad, 3 * T, s));
        HYMD_TRY(fft_c2r_z(pm->fft, grad, force_mesh, 3 * T, 1.f, s));
    }
    HYMD_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nv; ++j) moments[j] = ((double *)pm->host.ptr)[j];
    return HYMD_OK;
}

extern "C" int hymd_compute_field_force(hymd_pm *pm, const float *x, const int32_t *types, int64_t n,
                                        int32_t n_types, const float *force_mesh, float *force) {
    HYMD_REQUIRE(pm && force_mesh && force && (x || n == 0) && (types || n == 0), "null argument");
    HYMD_REQUIRE(n >= 0 && n_types >= 1 && n_types <= HYMD_MAX_TYPES, "bad sizes");
    HYMD_TRY(sort_particles(pm, x, types, nullptr, n));
    return readout_sorted(pm, 3, n_types, force_mesh, force, false);
}

extern "C" int hymd_update_field_force_q(hymd_pm *pm, const float *x, const float *charges, int64_t n,
                                         double sigma, double elec_factor, double self_energy,
                                         float *elec_forces, float *phi_q, float *psi, float *elec_field,
                                         double *energy_q) {
    HYMD_REQUIRE(pm && elec_forces && (n == 0 || (x && charges)) && n >= 0, "bad argument");
    const Mesh &m = pm->mesh;
    hipStream_t s = pm->stream;
    float2 *spec = (float2 *)pm->spec.ptr;
    float2 *out = (float2 *)pm->grad.ptr;
    float *rho = phi_q ? phi_q : (float *)pm->paint.ptr;
    HYMD_TRY(sort_particles(pm, x, nullptr, charges, n));
    HYMD_TRY(paint_sorted(pm, 1, nullptr, rho, nullptr));
    HYMD_TRY(fft_r2c_real(pm->fft, rho, spec, 1, s));
    PmeK p;
    p.sigma2 = sigma * sigma;
    p.inv_g = 1.0 / (double)m.g;
    p.four_pi_elec = 4.0 * M_PI * elec_factor;
    KGeom g = kgeom(pm);
    double *partials = energy_q ? (double *)pm->partials.ptr : nullptr;
    pme_kspace_kernel<<<pm->kblocks, kBlock, 0, s>>>(spec, out, psi ? 1 : 0, g, p, partials);
    HYMD_HIP(hipGetLastError());
    float *field = elec_field;
    if (!field) {
        HYMD_TRY(pm->rscratch.reserve(sizeof(float) * 3 * m.g));
        field = (float *)pm->rscratch.ptr;
    }
    HYMD_TRY(fft_inverse_xy(pm->fft, out, psi ? 4 : 3, s));
    HYMD_TRY(fft_c2r_z(pm->fft, out, field, 3, 1.f, s));
    if (psi) HYMD_TRY(fft_c2r_z(pm->fft, out + 3 * m.gc, psi, 1, 1.f, s));
    HYMD_TRY(readout_sorted(pm, 3, 1, field, elec_forces, true));
    if (energy_q) {
        double e = 0.0;
        HYMD_TRY(read_result(pm, pm->kblocks, 1, &e));
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
    DevBuf part, res;
    HYMD_TRY(part.reserve(sizeof(double) * nb));
    HYMD_TRY(res.reserve(sizeof(double)));
    self_energy_kernel<<<nb, kBlock, 0, s>>>(charges, n, (double *)part.ptr);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials((const double *)part.ptr, nb, 1, (double *)res.ptr, s));
    double q2 = 0.0;
    HYMD_HIP(hipMemcpyAsync(&q2, res.ptr, sizeof(double), hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    *out = elec_factor * std::sqrt(1.0 / (2.0 * M_PI * sigma * sigma)) * q2;
    return HYMD_OK;
}
