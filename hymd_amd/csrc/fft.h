// Batched mixed-radix (2,3,4,5,8) FFT for mesh edges 2^a 3^b 5^c, staged in LDS.
//
// Every 1-D transform is a self-sorting Stockham sequence: at a stage of radix
// R after sub-transforms of size p, butterfly i (of N/R) with k = i mod p reads
// x[i + j N/R], multiplies by w^(jk) (w = exp(-+2 pi i/(pR))), applies the
// R-point DFT and writes y[(i/p) p R + k + j p].  A workgroup holds W lines
// in two LDS buffers and ping-pongs between them, one barrier per stage.
//
// 3-D real transforms (unnormalised, numpy sign convention):
//   r2c: z rows -> packed complex FFT of length Mz/2 + split (post twiddles),
//        then in-place complex passes along y and x;
//   c2r: the same in reverse.
#pragma once
#include "common.h"

namespace hymd {

constexpr int kMaxStages = 12;
// Loads are issued kLoadBatch at a time per thread (all in flight before the
// first LDS store) so that every wave keeps several HBM requests outstanding.
constexpr int kLoadBatch = 8;

struct FftAxis {
    int n;        // complex line length
    int nstages;
    int ld;       // LDS row stride (complex elements)
    int radix[kMaxStages];
    int twoff[kMaxStages];  // offset of the stage's twiddles in the table
    FastDiv ndiv;               // n
    FastDiv pdiv[kMaxStages];   // p, the sub-transform size before the stage
    FastDiv nrdiv[kMaxStages];  // n / radix, the butterflies per line
};

struct PlaneDivs {
    FastDiv my, zh, nh, pairs;  // My, Mz/2+1, Mz/2, Mz/4+1
};

struct FftPlan {
    Mesh mesh;
    FftAxis ax, ay, azh;         // x, y full axes; z half-length packed axis
    int wx, wy, wz;              // lines per workgroup
    DevBuf tw_x, tw_y, tw_zh;    // forward stage twiddles (n-1 entries)
    DevBuf post;                 // exp(-2 pi i k / Mz), k = 0..Mz/2
    size_t lds_x, lds_y, lds_z;  // dynamic LDS bytes per kernel
    size_t lds_plane;            // y-z plane kernels (0: planes do not fit, separate passes)
    PlaneDivs pdiv;              // divisors of the plane kernels' index maps
};

int fft_plan_init(FftPlan &plan, const Mesh &mesh);

// 3-D real transforms of `batch` consecutive grids (unnormalised).  Where a
// My x (Mz/2+1) plane fits in LDS the z and y passes run as one plane kernel
// per (x-plane, grid), otherwise as separate line passes.
//   r2c:   fft_r2c_zy (z r2c + y) then fft_forward_x;  fft_r2c_real = both
//   c2r:   fft_inverse_x then fft_c2r_yz (y + z c2r, times `scale`);
//          fft_c2r = both.  The spectra passed in may be overwritten.
int fft_r2c_zy(const FftPlan &plan, const float *in, float2 *out, int batch, hipStream_t s);
int fft_forward_x(const FftPlan &plan, float2 *data, int batch, hipStream_t s);
int fft_r2c_real(const FftPlan &plan, const float *in, float2 *out, int batch, hipStream_t s);
int fft_inverse_x(const FftPlan &plan, float2 *data, int batch, hipStream_t s);
int fft_c2r_yz(const FftPlan &plan, float2 *in, float *out, int batch, float scale, hipStream_t s);
int fft_c2r(const FftPlan &plan, float2 *data, float *out, int batch, float scale, hipStream_t s);
// The 3T gradient grids out[3t + d] = c2r(-i k_d V_t) from the x-inverse
// transformed spectra in[t] = IFFT_x(-i k_x V_t) and in[T + t] = IFFT_x(V_t)
// (k_y = kfy n_y, k_z = kfz n_z, zero on Nyquist planes); needs plan.lds_plane.
int fft_c2r_yz_grad(const FftPlan &plan, const float2 *in, float *out, int T, double kfy, double kfz, float scale,
                    hipStream_t s);
// The same 3T gradient grids with the separate y and z line passes (meshes
// whose y-z planes do not fit in LDS): the y pass reads in[] and writes the
// 3T y-transformed spectra to tmp (3T half spectra), the z pass writes out.
int fft_c2r_lines_grad(const FftPlan &plan, const float2 *in, float2 *tmp, float *out, int T, double kfy, double kfz,
                       float scale, hipStream_t s);
// Whether fft_inplace<elems> can transform nlines lines of this axis with `threads` threads.
bool fft_inplace_fits(const FftAxis &ax, int nlines, int threads, int elems);

// ---------------------------------------------------------------------------
// device building blocks
// ---------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void small_dft(float2 (&u)[8], float sgn);

template <>
__device__ __forceinline__ void small_dft<2>(float2 (&u)[8], float) {
    float2 a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
}

template <>
__device__ __forceinline__ void small_dft<3>(float2 (&u)[8], float sgn) {
    const float h = 0.86602540378443864676f;  // sqrt(3)/2
    float2 t = cadd(u[1], u[2]);
    float2 m = make_float2(u[0].x - 0.5f * t.x, u[0].y - 0.5f * t.y);
    float2 s = cmul_i(csub(u[1], u[2]), sgn * h);
    u[0] = cadd(u[0], t);
    u[1] = cadd(m, s);
    u[2] = csub(m, s);
}

template <>
__device__ __forceinline__ void small_dft<4>(float2 (&u)[8], float sgn) {
    float2 t0 = cadd(u[0], u[2]), t1 = csub(u[0], u[2]);
    float2 t2 = cadd(u[1], u[3]), t3 = cmul_i(csub(u[1], u[3]), sgn);
    u[0] = cadd(t0, t2);
    u[1] = cadd(t1, t3);
    u[2] = csub(t0, t2);
    u[3] = csub(t1, t3);
}

template <>
__device__ __forceinline__ void small_dft<5>(float2 (&u)[8], float sgn) {
    const float c1 = 0.30901699437494742410f;   // cos(2pi/5)
    const float c2 = -0.80901699437494742410f;  // cos(4pi/5)
    const float s1 = 0.95105651629515357212f;   // sin(2pi/5)
    const float s2 = 0.58778525229247312917f;   // sin(4pi/5)
    float2 a1 = cadd(u[1], u[4]), b1 = csub(u[1], u[4]);
    float2 a2 = cadd(u[2], u[3]), b2 = csub(u[2], u[3]);
    float2 m1 = make_float2(u[0].x + c1 * a1.x + c2 * a2.x, u[0].y + c1 * a1.y + c2 * a2.y);
    float2 m2 = make_float2(u[0].x + c2 * a1.x + c1 * a2.x, u[0].y + c2 * a1.y + c1 * a2.y);
    float2 n1 = cmul_i(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y), sgn);
    float2 n2 = cmul_i(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y), sgn);
    u[0] = cadd(u[0], cadd(a1, a2));
    u[1] = cadd(m1, n1);
    u[4] = csub(m1, n1);
    u[2] = cadd(m2, n2);
    u[3] = csub(m2, n2);
}

template <>
__device__ __forceinline__ void small_dft<8>(float2 (&u)[8], float sgn) {
    const float r = 0.70710678118654752440f;
    float2 e[8] = {u[0], u[2], u[4], u[6], u[0], u[0], u[0], u[0]};
    float2 o[8] = {u[1], u[3], u[5], u[7], u[0], u[0], u[0], u[0]};
    small_dft<4>(e, sgn);
    small_dft<4>(o, sgn);
    // w^j o_j, w = exp(sgn 2 pi i / 8)
    float2 w1 = cscale(cadd(o[1], cmul_i(o[1], sgn)), r);
    float2 w2 = cmul_i(o[2], sgn);
    float2 w3 = cscale(csub(cmul_i(o[3], sgn), o[3]), r);
    u[0] = cadd(e[0], o[0]);
    u[4] = csub(e[0], o[0]);
    u[1] = cadd(e[1], w1);
    u[5] = csub(e[1], w1);
    u[2] = cadd(e[2], w2);
    u[6] = csub(e[2], w2);
    u[3] = cadd(e[3], w3);
    u[7] = csub(e[3], w3);
}

template <int R>
__device__ __forceinline__ void stockham_stage(const float2 *__restrict__ src, float2 *__restrict__ dst, int W,
                                               int ld, int N, int p, const FastDiv &fp, const FastDiv &fnr,
                                               const float2 *__restrict__ tw, float sgn) {
    const int NR = N / R;
    const int total = W * NR;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int l = fdiv(b, fnr);
        const int i = b - l * NR;
        const int q = fdiv(i, fp);
        const int k = i - q * p;
        const float2 *s = src + l * ld;
        float2 u[8];
#pragma unroll
        for (int j = 0; j < R; ++j) u[j] = s[i + j * NR];
        if (p > 1) {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                float2 w = tw[k * (R - 1) + j - 1];
                if (sgn > 0.f) w.y = -w.y;
                u[j] = cmul(u[j], w);
            }
        }
        small_dft<R>(u, sgn);
        float2 *d = dst + l * ld;
        const int ob = q * p * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) d[ob + j * p] = u[j];
    }
}

// Transforms W lines held in `a` (row stride ax.ld); returns the buffer that
// holds the result (a or b).  Caller must barrier after filling `a`.
__device__ __forceinline__ float2 *fft_lds(float2 *a, float2 *b, int W, const FftAxis &ax,
                                           const float2 *__restrict__ tw, float sgn) {
    int p = 1;
    for (int s = 0; s < ax.nstages; ++s) {
        const int R = ax.radix[s];
        const float2 *tws = tw + ax.twoff[s];
        const FastDiv &fp = ax.pdiv[s], &fnr = ax.nrdiv[s];
        switch (R) {
            case 2: stockham_stage<2>(a, b, W, ax.ld, ax.n, p, fp, fnr, tws, sgn); break;
            case 3: stockham_stage<3>(a, b, W, ax.ld, ax.n, p, fp, fnr, tws, sgn); break;
            case 4: stockham_stage<4>(a, b, W, ax.ld, ax.n, p, fp, fnr, tws, sgn); break;
            case 5: stockham_stage<5>(a, b, W, ax.ld, ax.n, p, fp, fnr, tws, sgn); break;
            default: stockham_stage<8>(a, b, W, ax.ld, ax.n, p, fp, fnr, tws, sgn); break;
        }
        __syncthreads();
        float2 *t = a;
        a = b;
        b = t;
        p *= R;
    }
    return a;
}

// ---------------------------------------------------------------------------
// In-place Stockham stage over lines with arbitrary strides (line l, element e
// at buf[l * sl + e * se]): every thread first gathers its butterflies (up to
// KB of radix R) into registers and transforms them, a barrier, then all
// write back.  Consecutive threads take consecutive lines.  Capacity:
// nlines * N <= kPlaneThreads * KB * R.
// ---------------------------------------------------------------------------
constexpr int kPlaneThreads = 512;
constexpr int kPlaneLoadBatch = 4;  // loads in flight per thread in the plane kernels (64-VGPR budget)
constexpr size_t kPlaneLdsBytes = 150 * 1024;  // of the 160 KiB a gfx950 workgroup may use
constexpr int kPlaneElems = 32;  // complex values per thread per stage (KB * R <= 32)

template <int R, int KB>
__device__ __forceinline__ void stockham_inplace(float2 *buf, int se, int sl, int nlines, const FastDiv &fl, int N,
                                                 int p, const FastDiv &fp, const float2 *__restrict__ tw, float sgn) {
    const int NR = N / R;
    const int total = nlines * NR;
    float2 u[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * blockDim.x;
        if (b >= total) continue;
        const int i = fdiv(b, fl), l = b - i * nlines, k = i - fdiv(i, fp) * p;
        const float2 *s = buf + l * sl;
#pragma unroll
        for (int j = 0; j < R; ++j) u[kb][j] = s[(i + j * NR) * se];
        if (p > 1) {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                float2 w = tw[k * (R - 1) + j - 1];
                if (sgn > 0.f) w.y = -w.y;
                u[kb][j] = cmul(u[kb][j], w);
            }
        }
        small_dft<R>(u[kb], sgn);
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * blockDim.x;
        if (b >= total) continue;
        const int i = fdiv(b, fl), l = b - i * nlines, q = fdiv(i, fp), k = i - q * p;
        float2 *d = buf + l * sl;
        const int ob = q * p * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) d[(ob + j * p) * se] = u[kb][j];
    }
    __syncthreads();
}

// In-place transform of nlines lines (strides as above) with at most ELEMS
// values per thread and stage; the caller barriers before (data ready) and
// may read the result right after.
template <int ELEMS = kPlaneElems>
__device__ __forceinline__ void fft_inplace(float2 *buf, int se, int sl, int nlines, const FastDiv &fl,
                                            const FftAxis &ax, const float2 *__restrict__ tw, float sgn) {
    int p = 1;
    for (int s = 0; s < ax.nstages; ++s) {
        const int R = ax.radix[s];
        const float2 *tws = tw + ax.twoff[s];
        const FastDiv &fp = ax.pdiv[s];
        switch (R) {
            case 2: stockham_inplace<2, ELEMS / 2>(buf, se, sl, nlines, fl, ax.n, p, fp, tws, sgn); break;
            case 3: stockham_inplace<3, ELEMS / 3>(buf, se, sl, nlines, fl, ax.n, p, fp, tws, sgn); break;
            case 4: stockham_inplace<4, ELEMS / 4>(buf, se, sl, nlines, fl, ax.n, p, fp, tws, sgn); break;
            case 5: stockham_inplace<5, ELEMS / 5>(buf, se, sl, nlines, fl, ax.n, p, fp, tws, sgn); break;
            default: stockham_inplace<8, ELEMS / 8>(buf, se, sl, nlines, fl, ax.n, p, fp, tws, sgn); break;
        }
        p *= R;
    }
}

// ---------------------------------------------------------------------------
// The same in-place stages with the transform length, the line count and the
// strides fixed at compile time: every index map is a multiply-shift or a
// shift, the strides fold into the LDS offsets, and each thread holds only
// the butterflies it owns.  The stage plan is the one fft_plan_init builds
// (StagePlan, from the same factorisation), so the twiddle tables are shared.
// ---------------------------------------------------------------------------
struct StagePlan {
    int ns;
    int r[kMaxStages], p[kMaxStages], twoff[kMaxStages];
};

// radices largest first: 8s, then one 4 / 4+4 / 2 for the power-of-two
// remainder, then 3s and 5s (ns = 0 if n is not 2^a 3^b 5^c)
constexpr StagePlan make_stage_plan(int n) {
    StagePlan sp{};
    int rs[kMaxStages] = {};
    int cnt = 0, a = 0;
    while (n % 2 == 0) {
        n /= 2;
        ++a;
    }
    int eights = a / 3;
    const int rem = a % 3;
    int tail[2] = {0, 0}, nt = 0;
    if (rem == 1 && eights > 0) {
        --eights;
        tail[nt++] = 4;
        tail[nt++] = 4;
    } else if (rem == 1) {
        tail[nt++] = 2;
    } else if (rem == 2) {
        tail[nt++] = 4;
    }
    for (int i = 0; i < eights; ++i) rs[cnt++] = 8;
    for (int i = 0; i < nt; ++i) rs[cnt++] = tail[i];
    while (n % 3 == 0) {
        n /= 3;
        rs[cnt++] = 3;
    }
    while (n % 5 == 0) {
        n /= 5;
        rs[cnt++] = 5;
    }
    if (n != 1 || cnt > kMaxStages) return sp;
    int p = 1, off = 0;
    for (int i = 0; i < cnt; ++i) {
        sp.r[i] = rs[i];
        sp.p[i] = p;
        sp.twoff[i] = off;
        off += p * (rs[i] - 1);
        p *= rs[i];
    }
    sp.ns = cnt;
    return sp;
}

template <int N>
struct CtPlan {
    static constexpr StagePlan v = make_stage_plan(N);
};

template <int R, int P, int N, int NLINES, int SE, int SL, int NT, int SGN>
__device__ __forceinline__ void stockham_inplace_ct(float2 *buf, const float2 *__restrict__ tw) {
    constexpr int NR = N / R, TOTAL = NLINES * NR, KB = (TOTAL + NT - 1) / NT;
    constexpr float sgn = (float)SGN;
    float2 u[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = b / NLINES, l = b - i * NLINES, k = i % P;
        const float2 *src = buf + l * SL + i * SE;
#pragma unroll
        for (int j = 0; j < R; ++j) u[kb][j] = src[j * NR * SE];
        if (P > 1) {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                float2 w = tw[k * (R - 1) + j - 1];
                if (SGN > 0) w.y = -w.y;
                u[kb][j] = cmul(u[kb][j], w);
            }
        }
        small_dft<R>(u[kb], sgn);
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = b / NLINES, l = b - i * NLINES, q = i / P, k = i - q * P;
        float2 *dst = buf + l * SL + (q * P * R + k) * SE;
#pragma unroll
        for (int j = 0; j < R; ++j) dst[j * P * SE] = u[kb][j];
    }
    __syncthreads();
}

// Stage 0 of fft_inplace_ct (P = 1, no twiddles) with its inputs from
// ld(line, element) (global memory) instead of buf: the load pass of the
// plane kernels and the first LDS round trip become one.  buf is only
// written, then a barrier; fft_inplace_ct<..., 1> runs the other stages.
// I_FAST: consecutive threads take consecutive butterflies of one line
// (the global reads of a line stored contiguously coalesce), else
// consecutive lines (lines stored side by side).
template <int N, int NLINES, int SE, int SL, int NT, int SGN, bool I_FAST = false, class Load>
__device__ __forceinline__ void fft_first_ct(float2 *buf, Load ld) {
    constexpr StagePlan sp = CtPlan<N>::v;
    static_assert(sp.ns >= 1, "no stage plan");
    constexpr int R = sp.r[0], NR = N / R, TOTAL = NLINES * NR, KB = (TOTAL + NT - 1) / NT;
    constexpr float sgn = (float)SGN;
    float2 u[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = I_FAST ? b % NR : b / NLINES, l = I_FAST ? b / NR : b - i * NLINES;
#pragma unroll
        for (int j = 0; j < R; ++j) u[kb][j] = ld(l, i + j * NR);
        small_dft<R>(u[kb], sgn);
        float2 *dst = buf + l * SL + i * R * SE;
#pragma unroll
        for (int j = 0; j < R; ++j) dst[j * SE] = u[kb][j];
    }
    __syncthreads();
}

// Stage 0 of fft_inplace_ct with its inputs formed by ld(line, element)
// from buf itself (a pass that would otherwise write buf first, such as the
// real-FFT split): every input is read before the barrier, every output
// written after it.
template <int N, int NLINES, int SE, int SL, int NT, int SGN, bool I_FAST, class Load>
__device__ __forceinline__ void fft_first_ct_inplace(float2 *buf, Load ld) {
    constexpr StagePlan sp = CtPlan<N>::v;
    static_assert(sp.ns >= 1, "no stage plan");
    constexpr int R = sp.r[0], NR = N / R, TOTAL = NLINES * NR, KB = (TOTAL + NT - 1) / NT;
    constexpr float sgn = (float)SGN;
    float2 u[KB][8];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = I_FAST ? b % NR : b / NLINES, l = I_FAST ? b / NR : b - i * NLINES;
#pragma unroll
        for (int j = 0; j < R; ++j) u[kb][j] = ld(l, i + j * NR);
        small_dft<R>(u[kb], sgn);
    }
    __syncthreads();
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = I_FAST ? b % NR : b / NLINES, l = I_FAST ? b / NR : b - i * NLINES;
        float2 *dst = buf + l * SL + i * R * SE;
#pragma unroll
        for (int j = 0; j < R; ++j) dst[j * SE] = u[kb][j];
    }
    __syncthreads();
}

template <int N, int NLINES, int SE, int SL, int NT, int SGN, int S = 0, int SEND = CtPlan<N>::v.ns>
__device__ __forceinline__ void fft_inplace_ct(float2 *buf, const float2 *__restrict__ tw) {
    constexpr StagePlan sp = CtPlan<N>::v;
    if constexpr (S < SEND) {
        stockham_inplace_ct<sp.r[S], sp.p[S], N, NLINES, SE, SL, NT, SGN>(buf, tw + sp.twoff[S]);
        fft_inplace_ct<N, NLINES, SE, SL, NT, SGN, S + 1, SEND>(buf, tw);
    }
}

// The last stage of fft_inplace_ct (q = 0: output element k + j P) with its
// outputs to st(line, element, value) (global memory) instead of buf: the
// store pass and the last LDS round trip become one.  buf is only read.
template <int N, int NLINES, int SE, int SL, int NT, int SGN, bool I_FAST, class Store>
__device__ __forceinline__ void fft_last_ct(const float2 *buf, const float2 *__restrict__ tw, Store st) {
    constexpr StagePlan sp = CtPlan<N>::v;
    static_assert(sp.ns >= 1, "no stage plan");
    constexpr int S = sp.ns - 1, R = sp.r[S], P = sp.p[S], NR = N / R, TOTAL = NLINES * NR,
                  KB = (TOTAL + NT - 1) / NT;
    static_assert(P == NR, "the last stage spans the line");
    constexpr float sgn = (float)SGN;
    const float2 *tws = tw + sp.twoff[S];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
        const int b = threadIdx.x + kb * NT;
        if (TOTAL % NT != 0 && b >= TOTAL) continue;
        const int i = I_FAST ? b % NR : b / NLINES, l = I_FAST ? b / NR : b - i * NLINES;
        const float2 *src = buf + l * SL + i * SE;
        float2 u[8];
#pragma unroll
        for (int j = 0; j < R; ++j) u[j] = src[j * NR * SE];
        if (P > 1) {
#pragma unroll
            for (int j = 1; j < R; ++j) {
                float2 w = tws[i * (R - 1) + j - 1];
                if (SGN > 0) w.y = -w.y;
                u[j] = cmul(u[j], w);
            }
        }
        small_dft<R>(u, sgn);
#pragma unroll
        for (int j = 0; j < R; ++j) st(l, i + j * P, u[j]);
    }
}

}  // namespace hymd
