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
    ax.ndiv = make_fastdiv((uint32_t)n);
    ax.nstages = (int)rs.size();
    ax.ld = n + pad;
    std::vector<float2> host(n > 1 ? n - 1 : 1, make_float2(0.f, 0.f));
    int p = 1, off = 0;
    for (int s = 0; s < ax.nstages; ++s) {
        int R = rs[s];
        ax.radix[s] = R;
        ax.twoff[s] = off;
        ax.pdiv[s] = make_fastdiv((uint32_t)p);
        ax.nrdiv[s] = make_fastdiv((uint32_t)(n / R));
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
    plan.pdiv = PlaneDivs{make_fastdiv(mesh.my), make_fastdiv(mesh.zh), make_fastdiv(mesh.mz / 2),
                          make_fastdiv(mesh.mz / 4 + 1)};
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
    // strided passes: twiddles, two line buffers
    plan.lds_x = sizeof(float2) * (size_t)(plan.ax.n + 2 * plan.wx * plan.ax.ld);
    plan.lds_y = sizeof(float2) * (size_t)(plan.ay.n + 2 * plan.wy * plan.ay.ld);
    plan.lds_z = sizeof(float2) * (size_t)(plan.azh.n + mesh.zh + 2 * plan.wz * plan.azh.ld);
    // y-z plane kernels: twiddles (y, z half), post twiddles, one My x zh plane;
    // every in-place stage must fit kPlaneElems values per thread
    plan.lds_plane = sizeof(float2) * ((size_t)mesh.my + (size_t)plan.azh.n + 2 * (size_t)mesh.zh +
                                       (size_t)mesh.my * mesh.zh);
    if (plan.lds_plane > kPlaneLdsBytes || !fft_inplace_fits(plan.ay, mesh.zh, kPlaneThreads, kPlaneElems) ||
        !fft_inplace_fits(plan.azh, mesh.my, kPlaneThreads, kPlaneElems))
        plan.lds_plane = 0;
    const size_t lds_max = 64 * 1024;  // the line kernels stay within the default dynamic-LDS limit
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
__global__ void __launch_bounds__(kBlock) r2c_z_kernel(const float *__restrict__ in_, float2 *__restrict__ out,
                                                        int64_t rows, int mz, int zh, FftAxis ax,
                                                        const float2 *__restrict__ tw_glob,
                                                        const float2 *__restrict__ post_glob, int W) {
    extern __shared__ float2 smem[];
    const int nh = ax.n;
    float2 *tw = smem;
    float2 *post = tw + nh;
    float2 *A = post + zh;
    float2 *B = A + W * ax.ld;
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_glob[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int64_t row0 = (int64_t)bx * W;
    const int64_t bofs = (int64_t)by * rows;
    // the W rows are contiguous: element idx of the block is complex z[idx % nh] of row idx / nh
    const int64_t nvalid = (int64_t)(rows - row0 < W ? rows - row0 : W) * nh;
    const int64_t gofs = (bofs + row0) * nh;  // in complex (pairs of reals) units
    const int total = W * nh;
    for (int i0 = threadIdx.x; i0 < total; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = make_float2(0.f, 0.f);
            if (idx < nvalid) v[j] = reinterpret_cast<const float2 *>(in_)[gofs + idx];
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < total) {
                const int l = idx / nh;
                A[l * ax.ld + (idx - l * nh)] = v[j];
            }
        }
    }
    __syncthreads();
    float2 *Z = fft_lds(A, B, W, ax, tw, -1.f);
    const int64_t nout = (int64_t)(rows - row0 < W ? rows - row0 : W) * zh;
    float2 *o = out + (bofs + row0) * zh;
    for (int idx = threadIdx.x; idx < nout; idx += blockDim.x) {
        const int l = idx / zh, k = idx - l * zh;
        const float2 *zr = Z + l * ax.ld;
        float2 zk = zr[k == nh ? 0 : k];
        float2 zc = cconj(zr[k == 0 ? 0 : nh - k]);
        float2 a = cscale(cadd(zk, zc), 0.5f);
        float2 b = cmul_i(csub(zk, zc), -0.5f);
        o[idx] = cadd(a, cmul(post[k], b));
    }
}

// r2c_z_kernel for a compile-time half length NH and W rows per workgroup
// (the plan's own stage plan and twiddles): the first stage reads the rows
// from global memory, the transform runs in place, then the split.
template <int NH, int W>
__global__ void __launch_bounds__(kBlock) r2c_z_ct_kernel(const float *__restrict__ in_, float2 *__restrict__ out,
                                                           int64_t rows, const float2 *__restrict__ tw_glob,
                                                           const float2 *__restrict__ post_glob) {
    constexpr int ZH = NH + 1;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *post = tw + NH;
    float2 *A = post + ZH;  // W rows at stride ZH
    for (int i = threadIdx.x; i < NH - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < ZH; i += blockDim.x) post[i] = post_glob[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int64_t row0 = (int64_t)bx * W;
    const int64_t bofs = (int64_t)by * rows;
    const int wv = (int)(rows - row0 < W ? rows - row0 : W);
    const float2 *src = reinterpret_cast<const float2 *>(in_) + (bofs + row0) * NH;  // rows of NH complex
    fft_first_ct<NH, W, 1, ZH, kBlock, -1, true>(A, [&](int l, int e) -> float2 {
        return l < wv ? src[l * NH + e] : make_float2(0.f, 0.f);
    });
    fft_inplace_ct<NH, W, 1, ZH, kBlock, -1, 1>(A, tw);
    float2 *o = out + (bofs + row0) * ZH;
    for (int idx = threadIdx.x; idx < wv * ZH; idx += blockDim.x) {
        const int l = idx / ZH, k = idx - l * ZH;
        const float2 *zr = A + l * ZH;
        float2 zk = zr[k == NH ? 0 : k];
        float2 zc = cconj(zr[k == 0 ? 0 : NH - k]);
        float2 a = cscale(cadd(zk, zc), 0.5f);
        float2 b = cmul_i(csub(zk, zc), -0.5f);
        o[idx] = cadd(a, cmul(post[k], b));
    }
}

// z pass, half spectrum -> real:  Z'[k] = (X[k] + X*[Nh-k]) + i e^{2 pi i k/Mz} (X[k] - X*[Nh-k]),
// inverse FFT of length Nh gives x[2n] + i x[2n+1] (unnormalised).
// With kfz != 0 the grids by = 3t + 2 are multiplied by -i k_z (k_z = kfz kz,
// zero on the Nyquist plane) as they are loaded: the z derivative of the
// gradient passes.
__global__ void __launch_bounds__(kBlock) c2r_z_kernel(const float2 *__restrict__ in, float *__restrict__ out,
                                                        int64_t rows, int mz, int zh, FftAxis ax,
                                                        const float2 *__restrict__ tw_glob,
                                                        const float2 *__restrict__ post_glob, float scale, int W,
                                                        float kfz) {
    extern __shared__ float2 smem[];
    const int nh = ax.n;
    float2 *tw = smem;
    float2 *post = tw + nh;
    float2 *A = post + zh;
    float2 *B = A + W * ax.ld;
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_glob[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int64_t row0 = (int64_t)bx * W;
    const int64_t bofs = (int64_t)by * rows;
    const int wv = (int)(rows - row0 < W ? rows - row0 : W);
    const float2 *src = in + (bofs + row0) * zh;   // W contiguous half-spectrum rows (ld = zh)
    const int nvalid = wv * zh, total = W * zh;
    const bool dz = kfz != 0.f && by % 3 == 2;
    for (int i0 = threadIdx.x; i0 < total; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = idx < nvalid ? src[idx] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < total) {
                if (dz) {   // -i k_z v = (k_z v.y, -k_z v.x)
                    const int kz = idx % zh;
                    const float k = (2 * kz == mz) ? 0.f : kfz * (float)kz;
                    v[j] = make_float2(k * v[j].y, -k * v[j].x);
                }
                A[idx] = v[j];
            }
        }
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
    float2 *dst = reinterpret_cast<float2 *>(out) + (bofs + row0) * nh;
    for (int idx = threadIdx.x; idx < wv * nh; idx += blockDim.x) {
        const int l = idx / nh, e = idx - l * nh;
        const float2 v = z[l * ax.ld + e];
        dst[idx] = make_float2(v.x * scale, v.y * scale);
    }
}

// c2r_z_kernel for a compile-time half length NH (rows of NH + 1 spectra)
// and W rows per workgroup (the plan's own stage plan and twiddles): the
// transform runs in place and its last stage writes the rows, scaled, to
// global memory.
template <int NH, int W>
__global__ void __launch_bounds__(kBlock) c2r_z_ct_kernel(const float2 *__restrict__ in, float *__restrict__ out,
                                                           int64_t rows, const float2 *__restrict__ tw_glob,
                                                           const float2 *__restrict__ post_glob, float scale,
                                                           float kfz) {
    constexpr int ZH = NH + 1, MZ = 2 * NH, NS = CtPlan<NH>::v.ns;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *post = tw + NH;
    float2 *A = post + ZH;  // W rows at stride ZH
    for (int i = threadIdx.x; i < NH - 1; i += blockDim.x) tw[i] = tw_glob[i];
    for (int i = threadIdx.x; i < ZH; i += blockDim.x) post[i] = post_glob[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int64_t row0 = (int64_t)bx * W;
    const int64_t bofs = (int64_t)by * rows;
    const int wv = (int)(rows - row0 < W ? rows - row0 : W);
    const float2 *src = in + (bofs + row0) * ZH;  // W contiguous half-spectrum rows
    const int nvalid = wv * ZH, total = W * ZH;
    const bool dz = kfz != 0.f && by % 3 == 2;
    for (int i0 = threadIdx.x; i0 < total; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = idx < nvalid ? src[idx] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < total) {
                if (dz) {  // -i k_z v = (k_z v.y, -k_z v.x)
                    const int kz = idx % ZH;
                    const float k = (2 * kz == MZ) ? 0.f : kfz * (float)kz;
                    v[j] = make_float2(k * v[j].y, -k * v[j].x);
                }
                A[idx] = v[j];
            }
        }
    }
    __syncthreads();
    // the pre-process formed element by element in the first stage, in place
    static_assert(NS >= 2, "first and last stages must differ");
    fft_first_ct_inplace<NH, W, 1, ZH, kBlock, 1, true>(A, [&](int l, int k) -> float2 {
        const float2 *xr = A + l * ZH;
        const float2 xk = xr[k];
        const float2 xc = cconj(xr[NH - k]);
        const float2 sm = cadd(xk, xc);
        const float2 df = cmul(cconj(post[k]), csub(xk, xc));
        return cadd(sm, cmul_i(df, 1.f));
    });
    float2 *dst = reinterpret_cast<float2 *>(out) + (bofs + row0) * NH;
    fft_inplace_ct<NH, W, 1, ZH, kBlock, 1, 1, NS - 1>(A, tw);
    fft_last_ct<NH, W, 1, ZH, kBlock, 1, true>(A, tw, [&](int l, int e, float2 v) {
        if (l < wv) dst[l * NH + e] = make_float2(v.x * scale, v.y * scale);
    });
}

// dynamic LDS bytes of r2c_z_ct_kernel and c2r_z_ct_kernel: twiddles, post
// factors and W rows
template <int NH, int W>
constexpr size_t z_ct_lds() { return sizeof(float2) * (size_t)(NH + (NH + 1) + W * (NH + 1)); }

// In-place complex pass along a strided axis.  Line l -> base (l / inner) *
// outer + (l % inner), element e at base + e * es.  2^wshift consecutive
// lines per workgroup (consecutive in memory except at `inner` boundaries).
__global__ void __launch_bounds__(kBlock) fft_strided_kernel(float2 *__restrict__ data, int64_t batch_stride,
                                                              int nlines, int inner, int64_t outer, int64_t es,
                                                              FftAxis ax, const float2 *__restrict__ tw_glob,
                                                              float sgn, int wshift) {
    extern __shared__ float2 smem[];
    const int n = ax.n;
    const int W = 1 << wshift;
    float2 *tw = smem;
    float2 *A = tw + n;
    float2 *B = A + W * ax.ld;
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    float2 *d = data + (int64_t)by * batch_stride;
    const int line0 = (int)bx * W;
    // element idx of the block is line idx & (W - 1), element idx >> wshift;
    // kBlock is a multiple of W, so a thread only ever touches line tid & (W-1)
    const int myl = (int)threadIdx.x & (W - 1);
    const int myline = line0 + myl;
    const int64_t mybase = myline < nlines ? (int64_t)(myline / inner) * outer + (myline % inner) : -1;
    const int nel = W * n;
    for (int i0 = threadIdx.x; i0 < nel; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = make_float2(0.f, 0.f);
            if (idx < nel && mybase >= 0) v[j] = d[mybase + (int64_t)(idx >> wshift) * es];
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < nel) A[myl * ax.ld + (idx >> wshift)] = v[j];
        }
    }
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) tw[i] = tw_glob[i];
    __syncthreads();
    float2 *r = fft_lds(A, B, W, ax, tw, sgn);
    if (mybase >= 0)
        for (int idx = threadIdx.x; idx < nel; idx += blockDim.x) d[mybase + (int64_t)(idx >> wshift) * es] = r[myl * ax.ld + (idx >> wshift)];
}

// fft_strided_kernel for a compile-time line length N, W lines per
// workgroup and sign SGN (the plan's own stage plan and twiddles): the first
// stage reads the lines from global memory, the middle stages run in place
// in LDS and the last stage writes the lines back.
template <int N, int W, int SGN, int NT>
__global__ void __launch_bounds__(NT) fft_strided_ct_kernel(float2 *__restrict__ data, int64_t batch_stride,
                                                                 int nlines, int inner, int64_t outer, int64_t es,
                                                                 const float2 *__restrict__ tw_glob) {
    constexpr int NS = CtPlan<N>::v.ns;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + N;  // W lines at stride N + 1
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    float2 *d = data + (int64_t)by * batch_stride;
    const int line0 = (int)bx * W;
    for (int i = threadIdx.x; i < N - 1; i += NT) tw[i] = tw_glob[i];
    auto base = [&](int l) -> int64_t {
        const int line = line0 + l;
        return line < nlines ? (int64_t)(line / inner) * outer + (line % inner) : -1;
    };
    fft_first_ct<N, W, 1, N + 1, NT, SGN>(A, [&](int l, int e) -> float2 {
        const int64_t b = base(l);
        return b >= 0 ? d[b + (int64_t)e * es] : make_float2(0.f, 0.f);
    });
    fft_inplace_ct<N, W, 1, N + 1, NT, SGN, 1, NS - 1>(A, tw);
    fft_last_ct<N, W, 1, N + 1, NT, SGN, false>(A, tw, [&](int l, int e, float2 v) {
        const int64_t b = base(l);
        if (b >= 0) d[b + (int64_t)e * es] = v;
    });
}

// Out-of-place inverse y pass of the gradient spectra: grid by = 3t + d of
// out[] from in[t] (d = 0: in[t] = IFFT_x(-i k_x V_t)) or in[T + t]
// (IFFT_x(V_t)), times -i k_y on load for d = 1 (k_y = kfy ny, zero on the
// Nyquist plane); d = 2 gets its -i k_z in the z pass.  Lines (x, kz) along y
// as fft_strided_kernel's y pass.
__global__ void __launch_bounds__(kBlock) fft_y_grad_kernel(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                             int64_t gc, int nlines, int inner, int64_t outer,
                                                             int64_t es, int T, FftAxis ax,
                                                             const float2 *__restrict__ tw_glob, float kfy,
                                                             int wshift) {
    extern __shared__ float2 smem[];
    const int n = ax.n;
    const int W = 1 << wshift;
    float2 *tw = smem;
    float2 *A = tw + n;
    float2 *B = A + W * ax.ld;
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int t = (int)by / 3, d = (int)by - 3 * t;
    const float2 *src = in + (int64_t)(d == 0 ? t : T + t) * gc;
    float2 *dst = out + (int64_t)by * gc;
    const int line0 = (int)bx * W;
    const int myl = (int)threadIdx.x & (W - 1);
    const int myline = line0 + myl;
    const int64_t mybase = myline < nlines ? (int64_t)(myline / inner) * outer + (myline % inner) : -1;
    const int nel = W * n;
    for (int i0 = threadIdx.x; i0 < nel; i0 += kBlock * kLoadBatch) {
        float2 v[kLoadBatch];
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            v[j] = make_float2(0.f, 0.f);
            if (idx < nel && mybase >= 0) v[j] = src[mybase + (int64_t)(idx >> wshift) * es];
        }
#pragma unroll
        for (int j = 0; j < kLoadBatch; ++j) {
            const int idx = i0 + j * kBlock;
            if (idx < nel) {
                const int e = idx >> wshift;
                if (d == 1) {   // -i k_y v = (k_y v.y, -k_y v.x)
                    const int ny = (2 * e < n) ? e : e - n;
                    const float k = (2 * e == n) ? 0.f : kfy * (float)ny;
                    v[j] = make_float2(k * v[j].y, -k * v[j].x);
                }
                A[myl * ax.ld + e] = v[j];
            }
        }
    }
    for (int i = threadIdx.x; i < n - 1; i += blockDim.x) tw[i] = tw_glob[i];
    __syncthreads();
    float2 *r = fft_lds(A, B, W, ax, tw, 1.f);
    if (mybase >= 0)
        for (int idx = threadIdx.x; idx < nel; idx += blockDim.x)
            dst[mybase + (int64_t)(idx >> wshift) * es] = r[myl * ax.ld + (idx >> wshift)];
}

// fft_y_grad_kernel for a compile-time line length N and W lines per
// workgroup (the plan's own stage plan and twiddles): the first stage reads
// the lines from global memory (with -i k_y for d = 1) and the last stage
// writes them, so the lines cross LDS in place once per middle stage only.
template <int N, int W>
__global__ void __launch_bounds__(kBlock) fft_y_grad_ct_kernel(const float2 *__restrict__ in,
                                                                float2 *__restrict__ out, int64_t gc, int nlines,
                                                                int inner, int64_t outer, int64_t es, int T,
                                                                const float2 *__restrict__ tw_glob, float kfy) {
    constexpr int NS = CtPlan<N>::v.ns;
    extern __shared__ float2 smem[];
    float2 *tw = smem;
    float2 *A = tw + N;  // W lines at stride N + 1
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const unsigned bx = flat % gridDim.x, by = flat / gridDim.x;
    const int t = (int)by / 3, d = (int)by - 3 * t;
    const float2 *src = in + (int64_t)(d == 0 ? t : T + t) * gc;
    float2 *dst = out + (int64_t)by * gc;
    const int line0 = (int)bx * W;
    for (int i = threadIdx.x; i < N - 1; i += kBlock) tw[i] = tw_glob[i];
    auto base = [&](int l) -> int64_t {
        const int line = line0 + l;
        return line < nlines ? (int64_t)(line / inner) * outer + (line % inner) : -1;
    };
    fft_first_ct<N, W, 1, N + 1, kBlock, 1>(A, [&](int l, int e) -> float2 {
        const int64_t b = base(l);
        if (b < 0) return make_float2(0.f, 0.f);
        float2 v = src[b + (int64_t)e * es];
        if (d == 1) {  // -i k_y v = (k_y v.y, -k_y v.x)
            const int ny = (2 * e < N) ? e : e - N;
            const float k = (2 * e == N) ? 0.f : kfy * (float)ny;
            v = make_float2(k * v.y, -k * v.x);
        }
        return v;
    });
    fft_inplace_ct<N, W, 1, N + 1, kBlock, 1, 1, NS - 1>(A, tw);
    fft_last_ct<N, W, 1, N + 1, kBlock, 1, false>(A, tw, [&](int l, int e, float2 v) {
        const int64_t b = base(l);
        if (b >= 0) dst[b + (int64_t)e * es] = v;
    });
}

// ---------------------------------------------------------------------------
// y-z plane kernels: one workgroup per (x-plane, grid) holds the plane's My x
// (Mz/2+1) half spectrum in LDS and runs the z and y passes in place, so a
// plane crosses HBM once per direction instead of twice.
// ---------------------------------------------------------------------------

// Plane geometry: any mesh whose plane fits (runtime sizes, FastDiv index
// maps, runtime stage plans) or a fixed (My, Mz) with everything known at
// compile time.
struct PlaneRt {
    static constexpr int nt = kPlaneThreads, wpe = 1;
    static constexpr bool kCt = false;
    int my, mz, nh, zh, npairs;
    PlaneDivs dv;
    FftAxis ay, azh;
    __device__ int div_nh(int i) const { return fdiv(i, dv.nh); }
    __device__ int div_zh(int i) const { return fdiv(i, dv.zh); }
    __device__ int div_pairs(int i) const { return fdiv(i, dv.pairs); }
    template <int SGN>
    __device__ void fft_z(float2 *D, const float2 *tw) const {
        fft_inplace(D, 1, zh, my, dv.my, azh, tw, (float)SGN);
    }
    template <int SGN>
    __device__ void fft_y(float2 *D, const float2 *tw) const {
        fft_inplace(D, zh, 1, zh, dv.zh, ay, tw, (float)SGN);
    }
};

// NT threads per workgroup and at least WPE waves per SIMD (the register
// budget): kPlaneThreads, or 1024 threads where that keeps more waves on a
// CU.  160 x 160 (a 104 KB plane: one workgroup per CU) runs 16 waves per CU
// instead of 8; 128 x 128 (66.5 KB: two per CU) runs 32 instead of 16 when
// held to 64 VGPRs (WPE 8; the c2r pass then keeps 12 B/lane in scratch and
// is still faster: 0.3637 -> 0.3616 ms per lipid/water outer step).
template <int MY, int MZ, int NT = kPlaneThreads, int WPE = 1>
struct PlaneCt {
    static constexpr int my = MY, mz = MZ, nh = MZ / 2, zh = MZ / 2 + 1, npairs = MZ / 4 + 1, nt = NT, wpe = WPE;
    __device__ int div_nh(int i) const { return i / nh; }
    __device__ int div_zh(int i) const { return i / zh; }
    __device__ int div_pairs(int i) const { return i / npairs; }
    // rows: lines of nh complex, element stride 1, line stride zh
    template <int SGN>
    __device__ void fft_z(float2 *D, const float2 *tw) const {
        fft_inplace_ct<nh, MY, 1, zh, NT, SGN>(D, tw);
    }
    // the same with the first stage's inputs from ld(y, k) (rows contiguous
    // in global memory: consecutive threads take consecutive k)
    template <int SGN, class Load>
    __device__ void fft_z_from(float2 *D, const float2 *tw, Load ld) const {
        fft_first_ct<nh, MY, 1, zh, NT, SGN, true>(D, ld);
        fft_inplace_ct<nh, MY, 1, zh, NT, SGN, 1>(D, tw);
    }
    // the same with the last stage's outputs to st(y, k, value)
    template <int SGN, class Store>
    __device__ void fft_z_to(float2 *D, const float2 *tw, Store st) const {
        fft_inplace_ct<nh, MY, 1, zh, NT, SGN, 0, CtPlan<nh>::v.ns - 1>(D, tw);
        fft_last_ct<nh, MY, 1, zh, NT, SGN, true>(D, tw, st);
    }
    // the same with the first stage's inputs formed in place by ld(y, k)
    // (the c2r pre-process) and the last stage's outputs to st(y, k, value)
    template <int SGN, class Load, class Store>
    __device__ void fft_z_from_to(float2 *D, const float2 *tw, Load ld, Store st) const {
        static_assert(CtPlan<nh>::v.ns >= 2, "first and last stages must differ");
        fft_first_ct_inplace<nh, MY, 1, zh, NT, SGN, true>(D, ld);
        fft_inplace_ct<nh, MY, 1, zh, NT, SGN, 1, CtPlan<nh>::v.ns - 1>(D, tw);
        fft_last_ct<nh, MY, 1, zh, NT, SGN, true>(D, tw, st);
    }
    // columns: lines of My, element stride zh, line stride 1
    template <int SGN>
    __device__ void fft_y(float2 *D, const float2 *tw) const {
        fft_inplace_ct<MY, zh, zh, 1, NT, SGN>(D, tw);
    }
    // the same with the first stage's inputs from ld(kz, y)
    template <int SGN, class Load>
    __device__ void fft_y_from(float2 *D, const float2 *tw, Load ld) const {
        fft_first_ct<MY, zh, zh, 1, NT, SGN>(D, ld);
        fft_inplace_ct<MY, zh, zh, 1, NT, SGN, 1>(D, tw);
    }
    // the same with the last stage's outputs to st(kz, y, value)
    template <int SGN, class Store>
    __device__ void fft_y_to(float2 *D, const float2 *tw, Store st) const {
        fft_inplace_ct<MY, zh, zh, 1, NT, SGN, 0, CtPlan<MY>::v.ns - 1>(D, tw);
        fft_last_ct<MY, zh, zh, 1, NT, SGN, false>(D, tw, st);
    }
    // the same with the first stage's inputs formed in place by ld(kz, y)
    // (the r2c split)
    template <int SGN, class Load, class Store>
    __device__ void fft_y_from_to(float2 *D, const float2 *tw, Load ld, Store st) const {
        static_assert(CtPlan<MY>::v.ns >= 2, "first and last stages must differ");
        fft_first_ct_inplace<MY, zh, zh, 1, NT, SGN, false>(D, ld);
        fft_inplace_ct<MY, zh, zh, 1, NT, SGN, 1, CtPlan<MY>::v.ns - 1>(D, tw);
        fft_last_ct<MY, zh, zh, 1, NT, SGN, false>(D, tw, st);
    }
    static constexpr bool kCt = true;
};

// forward: My rows of Mz reals -> packed complex FFT of Mz/2 per row + split
// (as r2c_z_kernel) -> FFT along y of the Mz/2+1 columns -> out plane
template <class G>
__global__ void __launch_bounds__(G::nt, G::wpe) r2c_zy_plane_kernel(const float *__restrict__ in,
                                                                      float2 *__restrict__ out, int mx, G g,
                                                                      const float2 *__restrict__ twy_g,
                                                                      const float2 *__restrict__ twz_g,
                                                                      const float2 *__restrict__ post_g) {
    extern __shared__ float2 smem[];
    const int my = g.my, nh = g.nh, zh = g.zh;
    float2 *twy = smem;
    float2 *twz = twy + my;
    float2 *post = twz + nh;
    float2 *D = post + zh;  // my x zh, row stride zh
    for (int i = threadIdx.x; i < my - 1; i += blockDim.x) twy[i] = twy_g[i];
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) twz[i] = twz_g[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_g[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    const int64_t plane = (int64_t)(flat / gridDim.x) * mx + (flat % gridDim.x);  // grid * mx + x
    // rows of Mz reals as nh complex, row stride zh in D
    const float2 *src = reinterpret_cast<const float2 *>(in) + plane * my * nh;
    if constexpr (G::kCt) {
        // the z pass's first stage reads the rows from global memory
        g.template fft_z_from<-1>(D, twz, [&](int y, int k) -> float2 { return src[y * nh + k]; });
    } else {
        const int nel = my * nh;
        for (int i0 = threadIdx.x; i0 < nel; i0 += G::nt * kPlaneLoadBatch) {
            float2 v[kPlaneLoadBatch];
#pragma unroll
            for (int j = 0; j < kPlaneLoadBatch; ++j) {
                const int idx = i0 + j * G::nt;
                v[j] = idx < nel ? src[idx] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < kPlaneLoadBatch; ++j) {
                const int idx = i0 + j * G::nt;
                if (idx < nel) {
                    const int y = g.div_nh(idx);
                    D[y * zh + (idx - y * nh)] = v[j];
                }
            }
        }
        __syncthreads();
        g.template fft_z<-1>(D, twz);
    }
    float2 *dst = out + plane * my * zh;
    if constexpr (G::kCt) {
        // the split X[m] = (Z[m] + Z*[nh-m])/2 + post[m] (Z[m] - Z*[nh-m])/(2i)
        // (Z indices mod nh) formed element by element in the y pass's first
        // stage, the last stage writing the plane to global memory
        g.template fft_y_from_to<-1>(
            D, twy,
            [&](int kz, int y) -> float2 {
                const float2 *row = D + y * zh;
                const float2 zm = row[kz == nh ? 0 : kz], zc = row[kz == 0 ? 0 : nh - kz];
                const float2 a = cscale(cadd(zm, cconj(zc)), 0.5f);
                const float2 b = cmul_i(csub(zm, cconj(zc)), -0.5f);
                return cadd(a, cmul(post[kz], b));
            },
            [&](int kz, int y, float2 v) { dst[y * zh + kz] = v; });
        return;
    }
    // split in place: the thread of pair (k, nh - k) writes X[k] and X[nh - k]
    // (k = 0 writes X[0] and X[nh]; k = nh/2 is its own pair)
    const int npairs = g.npairs;
    for (int idx = threadIdx.x; idx < my * npairs; idx += blockDim.x) {
        const int y = g.div_pairs(idx), k = idx - y * npairs;
        float2 *row = D + y * zh;
        const int kc = k == 0 ? 0 : nh - k;
        const float2 zk = row[k], zc = row[kc];
        // X[k] = (Z[k] + Z*[nh-k])/2 + post[k] (Z[k] - Z*[nh-k])/(2i), Z[nh] := Z[0]
        const float2 a1 = cscale(cadd(zk, cconj(zc)), 0.5f);
        const float2 b1 = cmul_i(csub(zk, cconj(zc)), -0.5f);
        const float2 xk = cadd(a1, cmul(post[k], b1));
        const int km = k == 0 ? nh : kc;  // the partner output index
        const float2 a2 = cscale(cadd(zc, cconj(zk)), 0.5f);
        const float2 b2 = cmul_i(csub(zc, cconj(zk)), -0.5f);
        const float2 xm = cadd(a2, cmul(post[km], b2));
        row[k] = xk;
        if (km != k) row[km] = xm;
    }
    __syncthreads();
    g.template fft_y<-1>(D, twy);
    for (int idx = threadIdx.x; idx < my * zh; idx += blockDim.x) dst[idx] = D[idx];
}

// inverse: FFT along y of the Mz/2+1 columns -> per row Z'[k] = (X[k] + X*[nh-k])
// + i e^{2 pi i k/Mz} (X[k] - X*[nh-k]) (as c2r_z_kernel) -> inverse FFT of
// Mz/2 -> My rows of Mz reals, times `scale`.
// GRAD: the 3T gradient grids of fft_c2r_yz_grad.  Logical block
// (t Mx + x) 3 + d (the two planes that read in[T + t] are neighbours, so the
// second read is served by L2); d = 0 transforms in[t], d = 1 multiplies
// in[T + t] by -i k_y before the y pass, d = 2 by -i k_z after it; out grid 3t + d.
struct GradK {
    int T;
    double kfy, kfz;  // 2 pi / L_y, 2 pi / L_z
};

// -i kd v for a derivative factor kd
__device__ __forceinline__ float2 mul_minus_i(float2 v, float kd) { return make_float2(kd * v.y, -kd * v.x); }

template <class G, bool GRAD>
__global__ void __launch_bounds__(G::nt, G::wpe) c2r_yz_plane_kernel(const float2 *__restrict__ in,
                                                                      float *__restrict__ out, int mx, G g,
                                                                      const float2 *__restrict__ twy_g,
                                                                      const float2 *__restrict__ twz_g,
                                                                      const float2 *__restrict__ post_g,
                                                                      float scale, GradK gk) {
    extern __shared__ float2 smem[];
    const int my = g.my, mz = g.mz, nh = g.nh, zh = g.zh;
    float2 *twy = smem;
    float2 *twz = twy + my;
    float2 *post = twz + nh;
    float2 *D = post + zh;  // my x zh, row stride zh
    for (int i = threadIdx.x; i < my - 1; i += blockDim.x) twy[i] = twy_g[i];
    for (int i = threadIdx.x; i < nh - 1; i += blockDim.x) twz[i] = twz_g[i];
    for (int i = threadIdx.x; i < zh; i += blockDim.x) post[i] = post_g[i];
    const unsigned flat = xcd_block(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x * gridDim.y);
    int64_t in_plane, out_plane;
    int d = 0;
    if (GRAD) {
        d = (int)(flat % 3u);
        const int rest = (int)(flat / 3u), x = rest % mx, t = rest / mx;
        in_plane = (int64_t)(d == 0 ? t : gk.T + t) * mx + x;
        out_plane = (int64_t)(3 * t + d) * mx + x;
    } else {
        in_plane = out_plane = (int64_t)(flat / gridDim.x) * mx + (flat % gridDim.x);
    }
    const float2 *src = in + in_plane * my * zh;
    if constexpr (G::kCt) {
        // the y pass's first stage reads the plane from global memory
        g.template fft_y_from<1>(D, twy, [&](int kz, int y) -> float2 {
            float2 w = src[y * zh + kz];
            if (GRAD && d == 1) {  // row y: k_y, zero on the Nyquist row
                const int ny = 2 * y < my ? y : y - my;
                w = mul_minus_i(w, 2 * y == my ? 0.f : (float)(gk.kfy * ny));
            }
            return w;
        });
    } else {
        const int nel = my * zh;
        for (int i0 = threadIdx.x; i0 < nel; i0 += G::nt * kPlaneLoadBatch) {
            float2 v[kPlaneLoadBatch];
#pragma unroll
            for (int j = 0; j < kPlaneLoadBatch; ++j) {
                const int idx = i0 + j * G::nt;
                v[j] = idx < nel ? src[idx] : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int j = 0; j < kPlaneLoadBatch; ++j) {
                const int idx = i0 + j * G::nt;
                if (idx < nel) {
                    float2 w = v[j];
                    if (GRAD && d == 1) {  // row y: k_y, zero on the Nyquist row
                        const int y = g.div_zh(idx);
                        const int ny = 2 * y < my ? y : y - my;
                        w = mul_minus_i(w, 2 * y == my ? 0.f : (float)(gk.kfy * ny));
                    }
                    D[idx] = w;
                }
            }
        }
        __syncthreads();
        g.template fft_y<1>(D, twy);
    }
    float2 *dst = reinterpret_cast<float2 *>(out) + out_plane * my * nh;
    if constexpr (G::kCt) {
        // the pre-process Z[m] = (X[m] + X*[nh-m]) + i conj(post[m]) (X[m] - X*[nh-m])
        // (m = 0 reads X[nh]; with -i k_z for d = 2) formed element by element in
        // the z pass's first stage, the last stage writing the rows to global memory
        g.template fft_z_from_to<1>(
            D, twz,
            [&](int y, int m) -> float2 {
                const float2 *row = D + y * zh;
                const int p = nh - m;
                float2 xm = row[m], xp = row[p];
                if (GRAD && d == 2) {  // column kz: k_z, zero on the Nyquist column
                    xm = mul_minus_i(xm, 2 * m == mz ? 0.f : (float)(gk.kfz * m));
                    xp = mul_minus_i(xp, 2 * p == mz ? 0.f : (float)(gk.kfz * p));
                }
                const float2 sm = cadd(xm, cconj(xp));
                const float2 df = cmul(cconj(post[m]), csub(xm, cconj(xp)));
                return cadd(sm, cmul_i(df, 1.f));
            },
            [&](int y, int k, float2 v) { dst[y * nh + k] = make_float2(v.x * scale, v.y * scale); });
        return;
    }
    // pre-process in place, pairs (k, nh - k) per thread; k = 0 reads X[nh]
    const int npairs = g.npairs;
    for (int idx = threadIdx.x; idx < my * npairs; idx += blockDim.x) {
        const int y = g.div_pairs(idx), k = idx - y * npairs;
        float2 *row = D + y * zh;
        const int kc = nh - k;  // k = 0 -> X[nh]
        float2 xk = row[k], xc = row[kc];
        if (GRAD && d == 2) {  // column kz: k_z, zero on the Nyquist column
            xk = mul_minus_i(xk, 2 * k == mz ? 0.f : (float)(gk.kfz * k));
            xc = mul_minus_i(xc, 2 * kc == mz ? 0.f : (float)(gk.kfz * kc));
        }
        const float2 s1 = cadd(xk, cconj(xc));
        const float2 d1 = cmul(cconj(post[k]), csub(xk, cconj(xc)));
        const float2 zk = cadd(s1, cmul_i(d1, 1.f));
        if (k != 0 && kc != k) {
            const float2 s2 = cadd(xc, cconj(xk));
            const float2 d2 = cmul(cconj(post[kc]), csub(xc, cconj(xk)));
            row[kc] = cadd(s2, cmul_i(d2, 1.f));
        }
        row[k] = zk;
    }
    __syncthreads();
    g.template fft_z<1>(D, twz);
    for (int idx = threadIdx.x; idx < my * nh; idx += blockDim.x) {
        const int y = g.div_nh(idx);
        const float2 v = D[y * zh + (idx - y * nh)];
        dst[idx] = make_float2(v.x * scale, v.y * scale);
    }
}

static int log2_exact(int w) {
    int s = 0;
    while ((1 << s) < w) ++s;
    return s;
}

static bool same_plan(const FftAxis &ax, const StagePlan &sp) {
    if (ax.nstages != sp.ns) return false;
    for (int i = 0; i < sp.ns; ++i)
        if (ax.radix[i] != sp.r[i] || ax.twoff[i] != sp.twoff[i]) return false;
    return true;
}

// The compile-time strided pass when the axis has N points, W lines per
// workgroup and the stage plan of CtPlan<N>; false if not.  256-point lines
// take the W of the caller (their kernels keep one LDS buffer, so 16 lines
// fit); the others the plan's.
template <int N, int W>
static bool strided_ct(const FftPlan &plan, float2 *data, int batch, int axis, float sgn, hipStream_t s, int *err) {
    const Mesh &m = plan.mesh;
    const FftAxis &ax = axis == 1 ? plan.ay : plan.ax;
    const int w = axis == 1 ? plan.wy : plan.wx;
    if (ax.n != N || (N != 256 && w != W) || !same_plan(ax, CtPlan<N>::v)) return false;
    const int nlines = axis == 1 ? m.mx * m.zh : m.my * m.zh;
    const int inner = axis == 1 ? m.zh : nlines;
    const int64_t outer = axis == 1 ? (int64_t)m.my * m.zh : 0;
    const int64_t es = axis == 1 ? m.zh : nlines;
    dim3 grid((unsigned)((nlines + W - 1) / W), (unsigned)batch);
    const size_t lds = sizeof(float2) * (N + W * (N + 1));
    const float2 *tw = (const float2 *)(axis == 1 ? plan.tw_y.ptr : plan.tw_x.ptr);
    // 256-point lines at 16 per workgroup hold 35 KB of LDS, four workgroups
    // per CU: 512 threads keep twice the waves in flight
    constexpr int NT = (N == 256 && W == 16) ? 2 * kBlock : kBlock;
    if (sgn < 0.f)
        fft_strided_ct_kernel<N, W, -1, NT><<<grid, NT, lds, s>>>(data, m.gc, nlines, inner, outer, es, tw);
    else
        fft_strided_ct_kernel<N, W, 1, NT><<<grid, NT, lds, s>>>(data, m.gc, nlines, inner, outer, es, tw);
    const hipError_t e = hipGetLastError();
    *err = HYMD_OK;
    if (e != hipSuccess) {
        set_error("fft_strided_ct_kernel<%d, %d> launch failed: %s", N, W, hipGetErrorString(e));
        *err = HYMD_ERR_HIP;
    }
    return true;
}

static int launch_strided(const FftPlan &plan, float2 *data, int batch, int axis, float sgn, hipStream_t s) {
    const Mesh &m = plan.mesh;
    int err = HYMD_OK;
    // 256-point lines: 16 lines per workgroup (128-byte row segments, four
    // workgroups per CU) for batches of 4 grids or more, else 8 (eight per CU)
    if ((batch >= 4 ? strided_ct<256, 16>(plan, data, batch, axis, sgn, s, &err)
                    : strided_ct<256, 8>(plan, data, batch, axis, sgn, s, &err)) ||
        strided_ct<160, 16>(plan, data, batch, axis, sgn, s, &err) ||
        strided_ct<128, 16>(plan, data, batch, axis, sgn, s, &err) ||
        strided_ct<64, 16>(plan, data, batch, axis, sgn, s, &err) ||
        strided_ct<48, 16>(plan, data, batch, axis, sgn, s, &err))
        return err;
    if (axis == 1) {  // y: lines (x, kz), stride zh
        const int nlines = m.mx * m.zh;
        dim3 grid((unsigned)((nlines + plan.wy - 1) / plan.wy), (unsigned)batch);
        fft_strided_kernel<<<grid, kBlock, plan.lds_y, s>>>(data, m.gc, nlines, m.zh, (int64_t)m.my * m.zh, m.zh,
                                                             plan.ay, (const float2 *)plan.tw_y.ptr, sgn,
                                                             log2_exact(plan.wy));
    } else {  // x: lines (y, kz) contiguous, stride my*zh
        const int nlines = m.my * m.zh;
        dim3 grid((unsigned)((nlines + plan.wx - 1) / plan.wx), (unsigned)batch);
        fft_strided_kernel<<<grid, kBlock, plan.lds_x, s>>>(data, m.gc, nlines, nlines, 0, nlines, plan.ax,
                                                             (const float2 *)plan.tw_x.ptr, sgn,
                                                             log2_exact(plan.wx));
    }
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// Opt in to more than 64 KiB of dynamic LDS for the plane kernels of one
// geometry (once per instantiation).
template <class G>
static int plane_lds_opt_in() {
    static bool done = false;
    if (!done) {
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&r2c_zy_plane_kernel<G>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPlaneLdsBytes));
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&c2r_yz_plane_kernel<G, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPlaneLdsBytes));
        HYMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&c2r_yz_plane_kernel<G, true>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kPlaneLdsBytes));
        done = true;
    }
    return HYMD_OK;
}

// Calls f(geometry) with the compile-time geometry of the plan's (My, Mz) when
// one is built for it, the runtime one otherwise.
template <class F>
static int with_plane_geom(const FftPlan &plan, F &&f) {
    const Mesh &m = plan.mesh;
#define HYMD_PLANE_CT(MY, MZ)                                                                         \
    if (m.my == MY && m.mz == MZ && same_plan(plan.ay, CtPlan<MY>::v) && same_plan(plan.azh, CtPlan<MZ / 2>::v)) \
        return f(PlaneCt<MY, MZ>{});
    HYMD_PLANE_CT(48, 48)
    HYMD_PLANE_CT(64, 64)
    HYMD_PLANE_CT(96, 96)
#undef HYMD_PLANE_CT
    if (m.my == 128 && m.mz == 128 && same_plan(plan.ay, CtPlan<128>::v) && same_plan(plan.azh, CtPlan<64>::v))
        return f(PlaneCt<128, 128, 1024, 8>{});
    if (m.my == 160 && m.mz == 160 && same_plan(plan.ay, CtPlan<160>::v) && same_plan(plan.azh, CtPlan<80>::v))
        return f(PlaneCt<160, 160, 1024>{});
    return f(PlaneRt{m.my, m.mz, m.mz / 2, m.zh, m.mz / 4 + 1, plan.pdiv, plan.ay, plan.azh});
}

int fft_forward_x(const FftPlan &plan, float2 *data, int batch, hipStream_t s) {
    return launch_strided(plan, data, batch, 0, -1.f, s);
}

int fft_inverse_x(const FftPlan &plan, float2 *data, int batch, hipStream_t s) {
    return launch_strided(plan, data, batch, 0, 1.f, s);
}

int fft_r2c_zy(const FftPlan &plan, const float *in, float2 *out, int batch, hipStream_t s) {
    const Mesh &m = plan.mesh;
    if (plan.lds_plane) {
        return with_plane_geom(plan, [&](auto g) -> int {
            using G = decltype(g);
            HYMD_TRY(plane_lds_opt_in<G>());
            dim3 grid((unsigned)m.mx, (unsigned)batch);
            r2c_zy_plane_kernel<G><<<grid, G::nt, plan.lds_plane, s>>>(
                in, out, m.mx, g, (const float2 *)plan.tw_y.ptr, (const float2 *)plan.tw_zh.ptr,
                (const float2 *)plan.post.ptr);
            HYMD_HIP(hipGetLastError());
            return HYMD_OK;
        });
    }
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 grid((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)batch);
    if (m.mz == 256 && plan.wz == 16 && plan.azh.ld == 129 && same_plan(plan.azh, CtPlan<128>::v))
        r2c_z_ct_kernel<128, 16><<<grid, kBlock, z_ct_lds<128, 16>(), s>>>(in, out, rows, (const float2 *)plan.tw_zh.ptr,
                                                                  (const float2 *)plan.post.ptr);
    else
        r2c_z_kernel<<<grid, kBlock, plan.lds_z, s>>>(in, out, rows, m.mz, m.zh, plan.azh,
                                                       (const float2 *)plan.tw_zh.ptr, (const float2 *)plan.post.ptr,
                                                       plan.wz);
    HYMD_HIP(hipGetLastError());
    return launch_strided(plan, out, batch, 1, -1.f, s);
}

int fft_r2c_real(const FftPlan &plan, const float *in, float2 *out, int batch, hipStream_t s) {
    HYMD_TRY(fft_r2c_zy(plan, in, out, batch, s));
    return fft_forward_x(plan, out, batch, s);
}

int fft_c2r_yz(const FftPlan &plan, float2 *in, float *out, int batch, float scale, hipStream_t s) {
    const Mesh &m = plan.mesh;
    if (plan.lds_plane) {
        return with_plane_geom(plan, [&](auto g) -> int {
            using G = decltype(g);
            HYMD_TRY(plane_lds_opt_in<G>());
            dim3 grid((unsigned)m.mx, (unsigned)batch);
            c2r_yz_plane_kernel<G, false><<<grid, G::nt, plan.lds_plane, s>>>(
                in, out, m.mx, g, (const float2 *)plan.tw_y.ptr, (const float2 *)plan.tw_zh.ptr,
                (const float2 *)plan.post.ptr, scale, GradK{});
            HYMD_HIP(hipGetLastError());
            return HYMD_OK;
        });
    }
    HYMD_TRY(launch_strided(plan, in, batch, 1, 1.f, s));
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 grid((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)batch);
    if (m.mz == 256 && plan.wz == 16 && plan.azh.ld == 129 && same_plan(plan.azh, CtPlan<128>::v))
        c2r_z_ct_kernel<128, 16><<<grid, kBlock, z_ct_lds<128, 16>(), s>>>(in, out, rows, (const float2 *)plan.tw_zh.ptr,
                                                                  (const float2 *)plan.post.ptr, scale, 0.f);
    else
        c2r_z_kernel<<<grid, kBlock, plan.lds_z, s>>>(in, out, rows, m.mz, m.zh, plan.azh,
                                                       (const float2 *)plan.tw_zh.ptr, (const float2 *)plan.post.ptr,
                                                       scale, plan.wz, 0.f);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

int fft_c2r_yz_grad(const FftPlan &plan, const float2 *in, float *out, int T, double kfy, double kfz, float scale,
                    hipStream_t s) {
    const Mesh &m = plan.mesh;
    if (!plan.lds_plane) {
        set_error("the gradient plane pass needs a y-z plane that fits in LDS");
        return HYMD_ERR_UNSUPPORTED;
    }
    return with_plane_geom(plan, [&](auto g) -> int {
        using G = decltype(g);
        HYMD_TRY(plane_lds_opt_in<G>());
        dim3 grid((unsigned)m.mx, (unsigned)(3 * T));
        c2r_yz_plane_kernel<G, true><<<grid, G::nt, plan.lds_plane, s>>>(
            in, out, m.mx, g, (const float2 *)plan.tw_y.ptr, (const float2 *)plan.tw_zh.ptr,
            (const float2 *)plan.post.ptr, scale, GradK{T, kfy, kfz});
        HYMD_HIP(hipGetLastError());
        return HYMD_OK;
    });
}

int fft_c2r_lines_grad(const FftPlan &plan, const float2 *in, float2 *tmp, float *out, int T, double kfy, double kfz,
                       float scale, hipStream_t s) {
    const Mesh &m = plan.mesh;
    HYMD_REQUIRE(kfy != 0.0 && kfz != 0.0, "the gradient passes need non-zero k factors");
    const int nlines = m.mx * m.zh;
    dim3 gy((unsigned)((nlines + plan.wy - 1) / plan.wy), (unsigned)(3 * T));
    if (m.my == 256 && plan.wy == 8 && same_plan(plan.ay, CtPlan<256>::v))
        fft_y_grad_ct_kernel<256, 8><<<gy, kBlock, sizeof(float2) * (256 + 8 * 257), s>>>(
            in, tmp, m.gc, nlines, m.zh, (int64_t)m.my * m.zh, m.zh, T, (const float2 *)plan.tw_y.ptr, (float)kfy);
    else
        fft_y_grad_kernel<<<gy, kBlock, plan.lds_y, s>>>(in, tmp, m.gc, nlines, m.zh, (int64_t)m.my * m.zh, m.zh, T,
                                                         plan.ay, (const float2 *)plan.tw_y.ptr, (float)kfy,
                                                         log2_exact(plan.wy));
    HYMD_HIP(hipGetLastError());
    const int64_t rows = (int64_t)m.mx * m.my;
    dim3 gz((unsigned)((rows + plan.wz - 1) / plan.wz), (unsigned)(3 * T));
    if (m.mz == 256 && plan.wz == 16 && plan.azh.ld == 129 && same_plan(plan.azh, CtPlan<128>::v))
        c2r_z_ct_kernel<128, 16><<<gz, kBlock, z_ct_lds<128, 16>(), s>>>(tmp, out, rows, (const float2 *)plan.tw_zh.ptr,
                                                                (const float2 *)plan.post.ptr, scale, (float)kfz);
    else
        c2r_z_kernel<<<gz, kBlock, plan.lds_z, s>>>(tmp, out, rows, m.mz, m.zh, plan.azh,
                                                    (const float2 *)plan.tw_zh.ptr, (const float2 *)plan.post.ptr,
                                                    scale, plan.wz, (float)kfz);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

bool fft_inplace_fits(const FftAxis &ax, int nlines, int threads, int elems) {
    for (int st = 0; st < ax.nstages; ++st) {
        const int R = ax.radix[st];
        if ((int64_t)nlines * ax.n > (int64_t)threads * (elems / R) * R) return false;
    }
    return true;
}

int fft_c2r(const FftPlan &plan, float2 *data, float *out, int batch, float scale, hipStream_t s) {
    HYMD_TRY(fft_inverse_x(plan, data, batch, s));
    return fft_c2r_yz(plan, data, out, batch, scale, s);
}

}  // namespace hymd
