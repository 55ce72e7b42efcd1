// rRESPA velocity-Verlet pieces in FP64 (integrator.py:9-73 composed as in
// main.py:802-887,1143-1169).  Forces arrive in float32; positions and
// velocities stay float64; a float32 copy of the positions feeds the paint
// and readout kernels.
#include <cmath>

#include "common.h"

namespace hymd {

__global__ void __launch_bounds__(kBlock) kick_drift_kernel(double *__restrict__ x, double *__restrict__ v,
                                                             float *__restrict__ x32, int64_t n,
                                                             const float *__restrict__ ff, double cf,
                                                             const float *__restrict__ fs, double cs, double dt,
                                                             double b0, double b1, double b2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n3 = 3 * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        const int d = (int)(i % 3);
        const double L = d == 0 ? b0 : (d == 1 ? b1 : b2);
        double vi = v[i];
        if (fs) vi = vi + cs * (double)fs[i];
        vi = vi + cf * (double)ff[i];
        const double xi = np_mod(x[i] + dt * vi, L);  // common.h
        v[i] = vi;
        x[i] = xi;
        if (x32) x32[i] = (float)xi;
    }
}

__global__ void __launch_bounds__(kBlock) kick_kernel(double *__restrict__ v, int64_t n, const float *__restrict__ f1,
                                                       double c1, const float *__restrict__ f2, double c2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n3 = 3 * n;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        double vi = v[i] + c1 * (double)f1[i];
        if (f2) vi = vi + c2 * (double)f2[i];
        v[i] = vi;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) ke_kernel(const T *__restrict__ v, int64_t n3,
                                                     double *__restrict__ partials) {
    double acc[1] = {0.0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride)
        acc[0] += (double)v[i] * (double)v[i];
    block_reduce_store<1>(acc, partials + blockIdx.x);
}

// Force-array helpers of the outer step with PME and dihedral terms (each
// element rounded as the equivalent torch expression rounds it):
//   mode 0: out = a + (float)b      mode 1: out += a
//   mode 2: out += (float)b         mode 3: out = a
__global__ void __launch_bounds__(kBlock) force_combine_kernel(float *__restrict__ out, const float *__restrict__ a,
                                                                const double *__restrict__ b, int64_t n3, int mode) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        switch (mode) {
            case 0: out[i] = a[i] + (float)b[i]; break;
            case 1: out[i] = out[i] + a[i]; break;
            case 2: out[i] = out[i] + (float)b[i]; break;
            default: out[i] = a[i]; break;
        }
    }
}

// float64 <-> float32 copies (round to nearest)
__global__ void __launch_bounds__(kBlock) convert_kernel(float *__restrict__ f32, double *__restrict__ f64, int64_t n,
                                                          int to_f32) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (to_f32) f32[i] = (float)f64[i];
        else f64[i] = (double)f32[i];
    }
}

int force_combine(float *out, const float *a, const double *b, int64_t n3, int mode, hipStream_t s) {
    if (n3 > 0) force_combine_kernel<<<grid_for(n3), kBlock, 0, s>>>(out, a, b, n3, mode);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int convert_f64_f32(float *f32, double *f64, int64_t n, bool to_f32, hipStream_t s) {
    if (n > 0) convert_kernel<<<grid_for(n), kBlock, 0, s>>>(f32, f64, n, to_f32 ? 1 : 0);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// x[r, d] *= f[d] for rows [r0, r1) of an (n, 3) array; x32 (optional)
// receives float(x) of those rows
template <typename T>
__global__ void __launch_bounds__(kBlock) scale_rows_kernel(T *__restrict__ x, float *__restrict__ x32, int64_t r0,
                                                             int64_t r1, double f0, double f1, double f2) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t n3 = 3 * (r1 - r0);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n3; i += stride) {
        const int64_t e = 3 * r0 + i;
        const int d = (int)(e % 3);
        const double f = d == 0 ? f0 : (d == 1 ? f1 : f2);
        const T v = (T)((double)x[e] * f);
        x[e] = v;
        if (x32) x32[e] = (float)v;
    }
}

}  // namespace hymd

using namespace hymd;

extern "C" int hymd_scale_positions(void *x, int32_t precision, float *x32, int64_t n, int64_t row_begin,
                                    int64_t row_end, const double factors[3], void *stream) {
    HYMD_REQUIRE(factors && n >= 0 && 0 <= row_begin && row_begin <= row_end && row_end <= n, "bad row range");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    HYMD_REQUIRE(row_end == row_begin || x, "NULL positions");
    const int64_t m = row_end - row_begin;
    hipStream_t s = (hipStream_t)stream;
    if (m > 0) {
        if (precision == HYMD_F64)
            scale_rows_kernel<double><<<grid_for(3 * m), kBlock, 0, s>>>((double *)x, x32, row_begin, row_end,
                                                                          factors[0], factors[1], factors[2]);
        else
            scale_rows_kernel<float><<<grid_for(3 * m), kBlock, 0, s>>>((float *)x, x32, row_begin, row_end,
                                                                         factors[0], factors[1], factors[2]);
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

extern "C" int hymd_respa_kick_drift(double *x, double *v, float *x32, int64_t n, const float *f_fast, double c_fast,
                                     const float *f_slow, double c_slow, double dt, const double box[3],
                                     void *stream) {
    HYMD_REQUIRE(n >= 0 && box && (n == 0 || (x && v && f_fast)), "bad argument");
    for (int d = 0; d < 3; ++d) HYMD_REQUIRE(box[d] > 0.0, "box edges must be positive");
    if (n > 0)
        kick_drift_kernel<<<grid_for(3 * n), kBlock, 0, (hipStream_t)stream>>>(x, v, x32, n, f_fast, c_fast, f_slow,
                                                                              c_slow, dt, box[0], box[1], box[2]);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

extern "C" int hymd_respa_kick(double *v, int64_t n, const float *f1, double c1, const float *f2, double c2,
                               void *stream) {
    HYMD_REQUIRE(n >= 0 && (n == 0 || (v && f1)), "bad argument");
    if (n > 0) kick_kernel<<<grid_for(3 * n), kBlock, 0, (hipStream_t)stream>>>(v, n, f1, c1, f2, c2);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

extern "C" int hymd_kinetic_energy(const void *v, int64_t n, int32_t precision, double mass, double *out,
                                   void *stream) {
    HYMD_REQUIRE(out && n >= 0 && (v || n == 0), "bad argument");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    hipStream_t s = (hipStream_t)stream;
    const int nb = grid_for(3 * n, kBlock, 1024);
    double *part = sync_scratch(nb + 1);  // nb partials, then the sum
    if (!part) return HYMD_ERR_NOMEM;
    if (precision == HYMD_F32)
        ke_kernel<float><<<nb, kBlock, 0, s>>>((const float *)v, 3 * n, part);
    else
        ke_kernel<double><<<nb, kBlock, 0, s>>>((const double *)v, 3 * n, part);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(reduce_partials(part, nb, 1, part + nb, s));
    double sum = 0.0;
    HYMD_HIP(hipMemcpyAsync(&sum, part + nb, sizeof(double), hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    *out = 0.5 * mass * sum;
    return HYMD_OK;
}
