// Shared helpers of the HIP hot path (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "hymd_amd.h"

namespace hymd {

void set_error(const char *fmt, ...);

#define HYMD_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            hymd::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
            return HYMD_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

#define HYMD_REQUIRE(cond, ...)          \
    do {                                 \
        if (!(cond)) {                   \
            hymd::set_error(__VA_ARGS__); \
            return HYMD_ERR_INVALID;     \
        }                                \
    } while (0)

#define HYMD_TRY(expr)           \
    do {                         \
        int s_ = (expr);         \
        if (s_ != HYMD_OK) return s_; \
    } while (0)

constexpr int kBlock = 256;

inline int grid_for(int64_t n, int block = kBlock, int cap = 1 << 20) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

// ---------------------------------------------------------------------------
// Cloud-in-cell window (pmesh CIC, node i at x = i*h).  Evaluated in float64
// with contraction disabled so that the CPU fixed-point replica computes the
// same bits: g = x*inv_h; i = floor(g); f = g - i; w0 = 1-f; w1 = f.
// ---------------------------------------------------------------------------
struct CicAxis {
    int i0, i1;
    double w0, w1;
};

// round(v) to int64, ties to even, the value __double2ll_rn gives.  For
// |v| < 2^51 (FAST) one FP64 addition does it: v + 1.5 * 2^52 lies in
// [2^52, 2^53), where the doubles are the integers, so the sum is rounded to
// the nearest integer (ties to even: 1.5 * 2^52 is even) and its low 52
// mantissa bits hold 2^51 + round(v); subtracting the bits of 1.5 * 2^52
// leaves round(v) in two's complement.  That is one FP64 add and one integer
// add instead of the six FP64 instructions of the general conversion.
// Callers pick FAST once per particle from a bound on |v| (fixed_fast).
constexpr double kFixedFastBound = 0x1p51;

template <bool FAST>
__device__ __forceinline__ long long fixed_round(double v) {
#pragma clang fp contract(off)
    if constexpr (FAST) return __double_as_longlong(v + 0x1.8p52) - 0x4338000000000000LL;
    else return __double2ll_rn(v);
}

__device__ __forceinline__ bool fixed_fast(double bound) { return fabs(bound) < kFixedFastBound; }

__device__ __forceinline__ CicAxis cic_axis(double x, double inv_h, int M) {
#pragma clang fp contract(off)
    CicAxis c;
    double g = x * inv_h;
    double fl = floor(g);
    double f = g - fl;
    // i0 = fl mod M, without 64-bit integer division: positions inside the
    // box take the first branch; the rest reduce exactly in float64 (integer
    // values below 2^53).  Non-finite input lands on node 0.
    int ii;
    if (fl >= 0.0 && fl < (double)M) {
        ii = (int)fl;
    } else {
        const double q = floor(fl / (double)M);
        ii = (int)(fl - (double)M * q);
        if (ii >= M) ii -= M;
        if (ii < 0) ii += M;
        if ((unsigned)ii >= (unsigned)M) ii = 0;
    }
    c.i0 = ii;
    c.i1 = (c.i0 + 1 == M) ? 0 : c.i0 + 1;
    c.w0 = 1.0 - f;
    c.w1 = f;
    return c;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  xcd_block maps the hardware block id to a logical one so that every
// XCD walks one contiguous range of logical blocks (a bijection on [0, nb)),
// which keeps neighbouring tiles, and the cache lines they share, in one L2.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb) {
    const unsigned per = nb / kXcds, rem = nb % kXcds, x = b % kXcds, q = b / kXcds;
    return x * per + (x < rem ? x : rem) + q;
}

// np.mod(x, L) for L > 0 as numpy computes it: fmod, negatives shifted up by
// L (so a tiny negative x gives L itself), zero returned as +0 (r + 0.0 turns
// -0 into +0 and leaves every other value as it is).  After a drift x lies
// in [-L, 2L), where x - L (exact by Sterbenz) and x + L are those values;
// anything else takes the general route.
__device__ __forceinline__ double np_mod(double x, double L) {
    double r;
    if (x >= 0.0 && x < L) {
        r = x;
    } else if (x >= L && x < 2.0 * L) {
        r = x - L;
    } else if (x < 0.0 && x >= -L) {
        r = x + L;
    } else {
        r = fmod(x, L);
        if (r != 0.0 && r < 0.0) r += L;
    }
    return r + 0.0;
}

// Rank of this lane's item among the items of the same bin, counted in
// counts[key] with atomics; key < 0 marks an inactive lane (no count, rank 0).
// Lanes of a wavefront that share a bin are aggregated: up to four rounds
// pick the first pending lane's bin and issue one atomic for all lanes in
// it; a round whose bin holds a single lane ends the aggregation (the input
// is not bin-ordered there), and the lanes left over use one atomic each.
// Spatially ordered input has few bins per wavefront, so its atomics (and
// their contention on one counter) drop by up to 64x.  Every lane of the
// wavefront must call it.
__device__ __forceinline__ int rank_in_bin(int key, int *counts) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(key >= 0);
    int rank = 0;
    for (int round = 0; round < 4 && pending; ++round) {
        const int leader = __ffsll((long long)pending) - 1;
        const int k = __shfl(key, leader, 64);
        const unsigned long long match = __ballot(key == k) & pending;
        int base = 0;
        if (lane == leader) base = atomicAdd(counts + k, __popcll(match));
        base = __shfl(base, leader, 64);
        if ((match >> lane) & 1ull) rank = base + __popcll(match & ((1ull << lane) - 1ull));
        pending &= ~match;
        if (__popcll(match) == 1) break;
    }
    if ((pending >> lane) & 1ull) rank = atomicAdd(counts + key, 1);
    return rank;
}

// rank_in_bin with every bin of the wavefront aggregated: the lanes are
// grouped first (one ballot per distinct bin, no memory traffic), then each
// group's lowest lane reserves the group's ranks with one atomic, all groups'
// atomics in flight together.  At most one atomic per distinct bin per
// wavefront, for ~10 instructions per distinct bin: the choice for key
// kernels whose waves do little else.  Every lane of the wavefront must call it.
__device__ __forceinline__ int rank_in_bin_all(int key, int *counts) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long pending = __ballot(key >= 0);
    int leader = lane, offset = 0, size = 0;
    while (pending) {  // uniform: one pass per distinct bin of the wavefront
        const int first = __ffsll((long long)pending) - 1;
        const int k = __builtin_amdgcn_readlane(key, first);
        const unsigned long long match = __ballot(key == k) & pending;
        if ((match >> lane) & 1ull) {
            leader = first;
            offset = __popcll(match & below);
            if (lane == first) size = __popcll(match);
        }
        pending &= ~match;
    }
    int base = 0;
    if (size) base = atomicAdd(counts + key, size);
    base = __shfl(base, leader, 64);
    return key >= 0 ? base + offset : 0;
}

// Division by a divisor fixed at launch, as a multiply-high and a shift
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up multiplier): with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1, n / d = (umulhi(n, m) + n) >> s for every
// 0 <= n < 2^31 (the sum cannot overflow since umulhi(n, m) <= n).  The FFT
// stages use it where an integer divide would cost ~20 instructions per lane.
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    const uint64_t m = (((1ull << s) - d) << 32) / d + 1;
    return FastDiv{d, (uint32_t)m, s};
}

__device__ __forceinline__ int fdiv(int n, const FastDiv &f) {
    return (int)((__umulhi((uint32_t)n, f.m) + (uint32_t)n) >> f.s);
}

// mesh geometry passed by value to kernels
struct Mesh {
    int mx, my, mz, zh;     // real edges, zh = mz/2+1
    int64_t g, gc;          // real cells, complex cells
    double inv_h[3];        // M_d / L_d
    double box[3];
};

// complex helpers (float2 = re, im)
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by +i (s>0) or -i (s<0)
__device__ __forceinline__ float2 cmul_i(float2 a, float s) { return make_float2(-s * a.y, s * a.x); }

// Deterministic block reduction of doubles (fixed order: a shuffle tree
// inside each 64-wide wavefront, then the wavefront sums in order).  All
// NT (the block size) threads of the block must call it.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out_block) {
    constexpr int kWaves = NT / 64;
    __shared__ double red[NV][kWaves];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < NV; ++j) {
        double s = v[j];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int j = 0; j < NV; ++j) {
            double t = 0.0;
            for (int w = 0; w < kWaves; ++w) t += red[j][w];
            out_block[j] = t;
        }
}

// Fixed-order sum over i = lane, lane + 64, ... < n of p[i * stride] (the
// order of a plain strided loop), with up to 16 loads in flight per lane.
__device__ __forceinline__ double lane_strided_sum(const double *__restrict__ p, int n, int64_t stride, int lane) {
    double t = 0.0;
    for (int b0 = lane; b0 < n; b0 += 64 * 16) {
        double x[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int b = b0 + 64 * q;
            x[q] = b < n ? p[(int64_t)b * stride] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) t += x[q];
    }
    return t;
}

// Reduce `nblocks` partial rows of NV doubles (row-major) in fixed order into
// out[NV] with a single block; launch with kBlock threads.
int reduce_partials(const double *partials, int nblocks, int nv, double *out, hipStream_t s);
// Clear nwords 32-bit words with a fill kernel (safe inside captured graphs).
int zero_words(void *ptr, int64_t nwords, hipStream_t s);

// nbatch independent reductions: batch r reduces nblocks[r] rows from row
// row0[r] into out[r * nv ..] (tables in device memory), same arithmetic.
int reduce_partials_batched(const double *partials, const int *nblocks, const int64_t *row0, int nv, double *out,
                            int nbatch, hipStream_t s);

// Exclusive scan of n int32 counts into out[0..n] (out[n] = total); tiles is
// scratch of scan_tiles(n) + 1 ints; cursor (optional) receives a copy of
// out[0..n-1]; zero_input clears the counts after reading them.
int64_t scan_tiles(int64_t n);
int exclusive_scan(int *in, int *out, int64_t n, int *tiles, int *cursor, bool zero_input, hipStream_t s);

// Scratch growing buffer owned by the library (per handle or per thread).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    uint64_t allocs = 0;  // successful allocations so far: a change means `ptr` moved
    int reserve(size_t need);
    void release();
    ~DevBuf() { release(); }
};

// Term arrays held by a persistent topology (device pointers).
struct TopoTerms {
    int64_t n, m2, m3;
    const int32_t *a2, *b2, *a3, *b3, *c3;
    const double *r0, *k2, *t0, *k3;
};
TopoTerms topology_terms(const hymd_topology *t);

// compute_field_force at the particle set sorted by the last update_field on
// this handle (same positions and types)
int field_readout_last(hymd_pm *pm, int n_types, const float *force_mesh, float *force);
// hymd_update_field_force_q with accumulate: elec_forces += the PME forces
// (the float32 sum of a separate addition pass; uncharged rows untouched)
int update_field_force_q_impl(hymd_pm *pm, const float *x, const float *charges, int64_t n, double sigma,
                              double elec_factor, double self_energy, float *elec_forces, float *phi_q, float *psi,
                              float *elec_field, double *energy_q, bool accumulate);

// Changes whenever a buffer of the handle is reallocated or its box is set:
// work captured against the handle (graphs hold its pointers and its mesh by
// value) is stale once this differs from its value at capture.
uint64_t pm_generation(const hymd_pm *pm);
// the handle's current box edges
void pm_box(const hymd_pm *pm, double box[3]);
// a new handle with src's mesh and box (own buffers and plans)
int pm_clone(const hymd_pm *src, int32_t max_types, void *stream, hymd_pm **out);

// CSVR thermostat (thermostat.hip): the groups of one call, in the
// reference's order, and the tables fixed for a particle set.
struct CsvrGroups {
    int n_groups;
    int n_com;                            // COM classes; class n_com = "none"
    int g2;                               // COM classes padded to the moments kernel's G
    int nb;                               // workgroups of the moments kernel
    int cls[HYMD_MAX_TYPES];              // COM class of group g, -1: all-particle update
    double n[HYMD_MAX_TYPES];             // particles in group g
    double R[HYMD_MAX_TYPES];             // N(0,1) draw of group g
    double S[HYMD_MAX_TYPES];             // chi^2(3 n_g - 1) draw of group g
    double half_mass, kt_per_particle, c; // m/2, 3/2 R_gas T_0, exp(-dt n_inner / tau)
};
struct CsvrSetup {
    CsvrGroups gp{};
    uint64_t tab = 0;  // class of each type, 4 bits per type
};
int csvr_setup(CsvrSetup *cs, int32_t n_types, const int32_t *type_group, int32_t n_groups, const int64_t *group_n,
               double mass, double gas_constant, double target_temperature, double c, int32_t remove_com);
// Thermostat v (FP64 (n,3)) in place on stream s with the draws of one call;
// with fkick, the velocities are first completed by the pending outer half
// kick v + ckick * fkick (float32 (n,3)), which the scale pass writes back.
int csvr_enqueue(const CsvrSetup &cs, double *v, const int32_t *types, int64_t n, const double *gaussians,
                 const double *chi_squares, const float *fkick, double ckick, double *work, double *ws,
                 hipStream_t s);
// v -= sum(v) / n_particles
int com_enqueue(double *v, int64_t n, int64_t n_particles, double *ws, hipStream_t s);
int64_t thermostat_workspace_doubles();

// Row-binning keys written by a producer kernel instead of the field path's
// own key pass: keys[i] = (mesh row of particle i's CIC base cell, its rank
// among that row's particles), counts[row] incremented once per particle.
struct RowKeyTarget {
    int2 *keys;
    int *counts;
    double inv_h[2];  // cells per nm along x, y
    int mx, my;
};
int field_key_target(hymd_pm *pm, int64_t n, RowKeyTarget *t);
// First row of the trailing run of one-particle molecules of a fused handle
// (their rows carry no bonded terms, so they may be permuted among themselves).
int64_t fused_single_tail(const hymd_fused *f);
// order[slot] = i for the n particles at x in mesh-row order (between updates).
int row_order(hymd_pm *pm, const float *x, int64_t n, int *order, hipStream_t s);
// hymd_update_field for force grids only, with the keys of the float32
// positions x already produced into field_key_target()'s arrays
int update_field_keyed(hymd_pm *pm, const float *x, const int32_t *types, int64_t n, int32_t n_types,
                       const double *coupling, const double *shift, double constant, const double *mass,
                       double sigma, float *force_mesh, double *energy, void *phi_fourier = nullptr);
// force arrays of the outer step: mode 0 out = a + (float)b, 1 out += a,
// 2 out += (float)b, 3 out = a; and float64 <-> float32 copies
int force_combine(float *out, const float *a, const double *b, int64_t n3, int mode, hipStream_t s);
int convert_f64_f32(float *f32, double *f64, int64_t n, bool to_f32, hipStream_t s);

// Dihedral outputs of a fused inner loop whose handle holds dihedral terms:
// the dihedral energy and clamp count of the last sub-step into energy[2]
// when energy != NULL.
struct FusedDihOut {
    double *energy;
};
// hymd_fused_inner that also produces the row keys of the final positions
// (key) and the dihedral outputs (dih)
int fused_inner_impl(hymd_fused *f, double *x, double *v, float *x32, float *f_fast, const float *f_slow, double h_in,
                     double h_out, double dt, int32_t n_inner, int32_t pending, const double box[3], double *energies,
                     double *virials, void *stream, const RowKeyTarget *key, const FusedDihOut *dih = nullptr);
// The term arrays of a persistent dihedral handle (device pointers).
struct DihTerms {
    int64_t m;
    const int32_t *a, *b, *c, *d, *dtype, *last;
    const float *coeff;
};
DihTerms dihedral_terms(const hymd_dihedrals *h);

// Pinned host staging for small results.
// Device scratch of at least n doubles for reductions that synchronise
// before returning: one grow-only buffer per device and host thread, never
// freed (no HIP call at process exit).  NULL on failure (message set).
double *sync_scratch(size_t n);

struct HostBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    void release();
    ~HostBuf() { release(); }
};

}  // namespace hymd
