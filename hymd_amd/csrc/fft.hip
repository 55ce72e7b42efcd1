// FFT kernels: packed r2c / c2r along z, in-place complex passes along y and x.
#include <cmath>
#include <vector>

#include "fft.h"

namespace hymd {

// n = 2^a 3^b 5^c -> radices, largest first (8s, then one 4/4+4/2 for the
// power-of-two remainder, then 3s and 5s).  Returns false for other n.
static bool factorize(int n, std::vector<int> &rs) {
    rs.clear();
    int a = 0;
    while (n % 2 == 0) { n /= 2; ++a; }
    int eights = a / 3, rem = a % 3;
    if (rem == 1 && eights > 0) { --eights; rs.push_back(4); rs.push_back(4); }
    else if (rem == 1) rs.push_back(2);
    else if (rem == 2) rs.push_back(4);
    for (int i = 0; i < eights; ++i) rs.insert(rs.begin(), 8);
    while (n % 3 == 0) { n /= 3; rs.push_back(3); }
    while (n % 5 == 0) { n /= 5; rs.push_back(5); }
    return n == 1 && (int)rs.size() <= kMaxStages;
}

static int axis_init(FftAxis &ax, DevBuf &tw, int n, int pad) {
    std::vector<int> rs;
    if (!factorize(n, rs)) {
        set_error("FFT length %d is not of the form 2^a 3^b 5^c", n);
        return HYMD_ERR_UNSUPPORTED;
    }
    ax.n = n;
    ax.nstages = (int)rs.size();
    ax.ld = n + pad;
    std::vector<float2> host(n > 1 ? n - 1 : 1, make_float2(0.f, 0.f));
    int p = 1, off = 0;
    for (int s = 0; s < ax.nstages; ++s) {
        int R = rs[s];
        ax.radix[s] = R;
        ax.twoff[s] = off;
        for (int k = 0; k < p; ++k)
            for (int j = 1; j < R; ++j) {
                double ang = -2.0 * M_PI * (double)(j * k) / (double)(p * R);
                host[off + k * (R - 1) + j - 1] = make_float2((float)cos(ang), (float)sin(ang));
            }
        off += p * (R - 1);
        p *= R;
    }
    HYMD_TRY(tw.reserve(sizeof(float2) * host.size()));
    HYMD_HIP(hipMemcpy(tw.ptr, host.data(), sizeof(float2) * host.size(), hipMemcpyHostToDevice));
    return HYMD_OK;
}

int fft_plan_init(FftPlan &plan, const Mesh &mesh) {
    plan.mesh = mesh;
    if (mesh.mz % 2 != 0) {
        set_error("mesh edge along z must be even (got %d)", mesh.mz);
        return HYMD_ERR_UNSUPPORTED;
    }
    HYMD_TRY(axis_init(plan.ax, plan.tw_x, mesh.mx, 1));
    HYMD_TRY(axis_init(plan.ay, plan.tw_y, mesh.my, 1));
    HYMD_TRY(axis_init(plan.azh, plan.tw_zh, mesh.mz / 2, 0));
    plan.azh.ld = mesh.zh;  // rows hold Mz/2+1 spectra in c2r
    std::vector<float2> post(mesh.zh);
    for (int k = 0; k < mesh.zh; ++k) {
        double ang = -2.0 * M_PI * (double)k / (double)mesh.mz;
        post[k] = make_float2((float)cos(ang), (float)sin(ang));
    }
    HYMD_TRY(plan.post.reserve(sizeof(float2) * mesh.zh));
    HYMD_HIP(hipMemcpy(plan.post.ptr, post.data(), sizeof(float2) * mesh.zh, hipMemcpyHostToDevice));
    auto lines_for = [](int n) { return n >= 512 ? 4 : (n >= 256 ? 8 : 16); };
    plan.wx = lines_for(mesh.mx);
    plan.wy = lines_for(mesh.my);
    plan.wz = std::max(1, 2048 / (mesh.mz / 2));
    plan.lds_x = sizeof(float2) * (size_t)(plan.ax.n + 2 * plan.wx * plan.ax.ld);
    plan.lds_y = sizeof(float2) * (size_t)(plan.ay.n + 2 * plan.wy * plan.ay.ld);
    plan.lds_z = sizeof(float2) * (size_t)(plan.azh.n + mesh.zh + 2 * plan.wz * plan.azh.ld);
    const size_t lds_max = 64 * 1024;  // default dynamic-LDS limit without an opt-in attribute
    if (plan.lds_x > lds_max || plan.lds_y > lds_max || plan.lds_z > lds_max) {
        set_error("mesh too large for the LDS-staged FFT");
        return HYMD_ERR_UNSUPPORTED;
    }
    return HYMD_OK;
}

// ---------------------------------------------------------------------------
// z pass, real -> half spectrum.  One workgroup: W rows of Mz reals, viewed as
// Mz/2 complex z[n] = x[2n] + i x[2n+1]; FFT; split
//   X[k] = (Z[k] + Z*[Nh-k])/2 + e^{-2 pi i k/Mz} (Z[k] - Z*[Nh-k])/(2i).
// ---------------------------------------------------------------------------
template <bool FIXED>
__global__ void __launch_bounds__(kBlock) r2c_z_kernel(const void *__restrict__ in_, float2 *__restrict__ out,
                                                        int64_t rows, int mz, int zh, FftAxis ax,
                                                        const float2 *__restrict__ tw_glob,
                                                        const float2 *__restrict__ post_glob, double scale, int W,
                                                        int clear) {
    extern __shared__ float2 smem[];
    const int nh = ax.n;
    float2 *tw = smem;
    float2 *post = tw + nh;
    float2 *A = post + zh;
    float2 *B = A + W * ax.ld;
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_glob[i];
    const int64_t row0 = (int64_t)blockIdx.x * W;
    const int64_t bofs = (int64_t)blockIdx.y * rows;
    for (int idx = threadIdx.x; idx < W * nh; idx += blockDim.x) {
        const int l = idx / nh, e = idx - l * nh;
        const int64_t row = row0 + l;
        float2 z = make_float2(0.f, 0.f);
        if (row < rows) {
            const int64_t off = (bofs + row) * mz + 2 * e;
            if (FIXED) {
                long long *p = (long long *)in_ + off;
                long long q0 = p[0], q1 = p[1];
                z = make_float2((float)((double)q0 * scale), (float)((double)q1 * scale));
                if (clear) { p[0] = 0; p[1] = 0; }
            } else {
                z = *reinterpret_cast<const float2 *>((const float *)in_ + off);
            }
        }
        A[l * ax.ld + e] = z;
    }
    __syncthreads();
    float2 *Z = fft_lds(A, B, W, ax, tw, -1.f);
    for (int idx = threadIdx.x; idx < W * zh; idx += blockDim.x) {
        const int l = idx / zh, k = idx - l * zh;
        const int64_t row = row0 + l;
        if (row >= rows) continue;
        const float2 *zr = Z + l * ax.ld;
        float2 zk = zr[k == nh ? 0 : k];
        float2 zc = cconj(zr[k == 0 ? 0 : nh - k]);
        float2 a = cscale(cadd(zk, zc), 0.5f);
        float2 b = cmul_i(csub(zk, zc), -0.5f);
        out[(bofs + row) * zh + k] = cadd(a, cmul(post[k], b));
    }
}

// z pass, half spectrum -> real:  Z'[k] = (X[k] + X*[Nh-k]) + i e^{2 pi i k/Mz} (X[k] - X*[Nh-k]),
// inverse FFT of length Nh gives x[2n] + i x[2n+1] (unnormalised).
__global__ void __launch_bounds__(kBlock) c2r_z_kernel(const float2 *__restrict__ in, float *__restrict__ out,
                                                        int64_t rows, int mz, int zh, FftAxis ax,
                                                        const float2 *__restrict__ tw_glob,
                                                        const float2 *__restrict__ post_glob, float scale, int W) {
    extern __shared__ float2 smem[];
    const int nh = ax.n;
    float2 *tw = smem;
    float2 *post = tw + nh;
    float2 *A = post + zh;
    float2 *B = A + W * ax.ld;
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_glob[i];
    const int64_t row0 = (int64_t)blockIdx.x * W;
    const int64_t bofs = (int64_t)blockIdx.y * rows;
    for (int idx = threadIdx.x; idx < W * zh; idx += blockDim.x) {
        const int l = idx / zh, k = idx - l * zh;
        const int64_t row = row0 + l;
        A[l * ax.ld + k] = (row < rows) ? in[(bofs + row) * zh + k] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < W * nh; idx += blockDim.x) {
        const int l = idx / nh, k = idx - l * nh;
        const float2 *xr = A + l * ax.ld;
        float2 xk = xr[k];
        float2 xc = cconj(xr[nh - k]);
        float2 s = cadd(xk, xc);
        float2 d = cmul(cconj(post[k]), csub(xk, xc));
        B[l * ax.ld + k] = cadd(s, cmul_i(d, 1.f));
    }
    __syncthreads();
    float2 *z = fft_lds(B, A, W, ax, tw, 1.f);
    for (int idx = threadIdx.x; idx < W * nh; idx += blockDim.x) {
        const int l = idx / nh, e = idx - l * nh;
        const int64_t row = row0 + l;
        if (row >= rows) continue;
        float2 v = z[l * ax.ld + e];
        *reinterpret_cast<float2 *>(out + (bofs + row) * mz + 2 * e) = make_float2(v.x * scale, v.y * scale);
    }
}

// In-place complex pass along a strided axis.  Line l -> base (l / inner) *
// outer + (l % inner), element e at base + e * es.  W consecutive lines per
// workgroup (consecutive in memory except at `inner` boundaries).
__global__ void __launch_bounds__(kBlock) fft_strided_kernel(float2 *__restrict__ data, int64_t batch_stride,
                                                              int64_t nlines, int64_t inner, int64_t outer,
                                                              int64_t es, FftAxis ax,
                                                              const float2 *__restrict__ tw_glob, float sgn, int W) {
    extern __shared__ float2 smem[];
    const int n = ax.n;
    float2 *tw = smem;
    float2 *A = tw + n;
    float2 *B = A + W * ax.ld;
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) tw[i] = tw_glob[i];
    float2 *d = data + (int64_t)blockIdx.y * batch_stride;
    const int64_t line0 = (int64_t)blockIdx.x * W;
    for (int idx = threadIdx.x; idx < W * n; idx += blockDim.x) {
        const int e = idx / W, l = idx - e * W;
        const int64_t line = line0 + l;
        float2 v = make_float2(0.f, 0.f);
        if (line < nlines) v = d[(line / inner) * outer + (line % inner) + e * es];
        A[l * ax.ld + e] = v;
    }
    __syncthreads();
    float2 *r = fft_lds(A, B, W, ax, tw, sgn);
    for (int idx = threadIdx.x; idx < W * n; idx += blockDim.x) {
        const int e = idx / W, l = idx - e * W;
        const int64_t line = line0 + l;
        if (line < nlines) d[(line / inner) * outer + (line % inner) + e * es] = r[l * ax.ld + e];
    }
}

static int launch_strided(const FftPlan &plan, float2 *data, int batch, int axis, float sgn, hipStream_t s) {
    const Mesh &m = plan.mesh;
    if (axis == 1) {  // y: lines (x, kz), stride zh
        const int64_t nlines = (int64_t)m.mx * m.zh;
        dim3 grid((unsigned)((nlines + plan.wy - 1) / plan.wy), (unsigned)batch);
        fft_strided_kernel<<<grid, kBlock, plan.lds_y, s>>>(data, m.gc, nlines, m.zh, (int64_t)m.my * m.zh, m.zh,
                                                             plan.ay, (const float2 *)plan.tw_y.ptr, sgn, plan.wy);
    } else {  // x: lines (y, kz) contiguous, stride my*zh
        const int64_t nlines = (int64_t)m.my * m.zh;
        dim3 grid((unsigned)((nlines + plan.wx - 1) / plan.wx), (unsigned)batch);
        fft_strided_kernel<<<grid, kBlock, plan.lds_x, s>>>(data, m.gc, nlines, nlines, 0, nlines, plan.ax,
                                                             (const float2 *)plan.tw_x.ptr, sgn, plan.wx);
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int fft_forward_xy(const FftPlan &plan, float2 *data, int batch, hipStream_t s) {
    HYMD_TRY(launch_strided(plan, data, batch, 1, -1.f, s));
    HYMD_TRY(launch_strided(plan, data, batch, 0, -1.f, s));
    return HYMD_OK;
}

int fft_inverse_xy(const FftPlan &plan, float2 *data, int batch, hipStream_t s) {
    HYMD_TRY(launch_strided(plan, data, batch, 0, 1.f, s));
    HYMD_TRY(launch_strided(plan, data, batch, 1, 1.f, s));
    return HYMD_OK;
}

int fft_r2c_fixed(const FftPlan &plan, int64_t *in, float2 *out, int batch, double scale, bool clear,
                  hipStream_t s) {
    const Mesh &m = plan.mesh;
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 grid((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)batch);
    r2c_z_kernel<true><<<grid, kBlock, plan.lds_z, s>>>(in, out, rows, m.mz, m.zh, plan.azh,
                                                         (const float2 *)plan.tw_zh.ptr,
                                                         (const float2 *)plan.post.ptr, scale, plan.wz,
                                                         clear ? 1 : 0);
    HYMD_HIP(hipGetLastError());
    return fft_forward_xy(plan, out, batch, s);
}

int fft_r2c_real(const FftPlan &plan, const float *in, float2 *out, int batch, hipStream_t s) {
    const Mesh &m = plan.mesh;
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 grid((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)batch);
    r2c_z_kernel<false><<<grid, kBlock, plan.lds_z, s>>>(in, out, rows, m.mz, m.zh, plan.azh,
                                                          (const float2 *)plan.tw_zh.ptr,
                                                          (const float2 *)plan.post.ptr, 1.0, plan.wz, 0);
    HYMD_HIP(hipGetLastError());
    return fft_forward_xy(plan, out, batch, s);
}

int fft_c2r_z(const FftPlan &plan, const float2 *in, float *out, int batch, float scale, hipStream_t s) {
    const Mesh &m = plan.mesh;
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 grid((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)batch);
    c2r_z_kernel<<<grid, kBlock, plan.lds_z, s>>>(in, out, rows, m.mz, m.zh, plan.azh,
                                                   (const float2 *)plan.tw_zh.ptr, (const float2 *)plan.post.ptr,
                                                   scale, plan.wz);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

}  // namespace hymd
