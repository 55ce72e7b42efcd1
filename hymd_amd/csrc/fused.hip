// Molecule-blocked fused rRESPA inner loop (main.py:830-887).
//
// Bonded terms couple particles of one molecule only, and the particle-field
// force is constant during the inner steps.  When every molecule is a
// contiguous index range, whole molecules are packed into blocks of at most
// kBlock * PPT particles and one workgroup advances its block through all
// respa_inner sub-steps: positions live in LDS, velocities and forces in
// registers, and each sub-step evaluates every bond / angle / dihedral term
// once into an LDS slot, after which each particle sums the slots of its
// terms in term order (bonds, then angles, then dihedrals), as the
// per-particle gather kernels do.  HBM is touched once per outer step: x, v,
// F_fast, F_slow in; x, x32, v, F_fast out.  The dipoles and transfer
// matrices of a protein's dihedrals at the final positions (main.py:863-881
// reconstructs them on the last inner step) come from hymd_dihedrals_dipoles
// after the launch.
//
// Kick order per outer step, identical to the unfused kernels:
//   [pending outer half kick of the previous step] + outer half kick (F_slow)
//   inner i: first half kick (F_fast), drift, wrap, F_fast(x), second half kick
// where the second half of sub-step i and the first half of i+1 are separate
// FP64 additions of the same force.
#include <algorithm>
#include <numeric>
#include <type_traits>
#include <vector>

#include "bonded.h"

namespace hymd {


struct FusedArgs {
    double *x, *v;
    float *x32, *f_fast;
    const float *f_slow;
    const int *blk_start;        // nb+1 global particle index
    const int *blk_bond_off;     // nb+1 into bond_ab / bond_par
    const uint32_t *bond_ab;     // a | b << 16 (block-local)
    const double2 *bond_par;     // r0, k
    const int *blk_ang_off;      // nb+1
    const uint32_t *ang_abc;     // a | b << 10 | c << 20 (block-local, < 1024)
    const double2 *ang_par;      // theta0, k
    const int *csr_off;          // N+1 (global)
    const uint32_t *csr_ent;     // (term << 4) | (kind << 2) | role, kind 0 bond, 1 angle, 2 dihedral;
                                 // term block-local
    // dihedrals (kernel variant DIH): per block the terms' block-local
    // indices (a | b << 10 | c << 20, d | dtype << 10 | last << 12), their
    // 30 coefficients and their cosine-series constants (kDihSeries per
    // term, constant-major within the block: constant k of term j of a block
    // with nd terms from offset d0 at dih_ser[kDihSeries d0 + k nd + j])
    const int *blk_dih_off;      // nb+1
    const uint2 *dih_idx;
    const float *dih_coeff;
    const double *dih_ser;
    double h_in, h_out, dt;
    double box[3];
    int n_inner;
    int pending;                 // apply the previous step's outer half kick first
    double *partials;            // per block E_bond, pr_bond[3], E_angle, pr_angle[3] (+ E_dihedral,
                                 // clamps with DIH), or null
    RowKeyTarget key;            // row-binning keys of the final positions (key.keys == null: none)
};

// Dynamic LDS of one block: positions (3 doubles per particle), bond slots
// (3 doubles), angle slots (6 doubles), dihedral slots (12 doubles),
// per-particle entry offsets (int) and entries (2 per bond, 3 per angle, 4
// per dihedral).
// Budgets: 256 particles per block (one per thread) in 26 KB, six blocks per
// CU; molecules above 256 beads use 512-particle blocks (two per thread) in
// 52 KB, three per CU.  With dihedrals (an FP64 term of ~200 registers' worth
// of state, so fewer waves anyway) 40 KB and 80 KB.
constexpr size_t kFusedLds1 = 26 * 1024;
constexpr size_t kFusedLds2 = 52 * 1024;
constexpr size_t kFusedLdsDih1 = 40 * 1024;
constexpr size_t kFusedLdsDih2 = 80 * 1024;
inline size_t fused_lds_bytes(int particles, int bonds, int angles, int dihedrals) {
    return sizeof(double) * (3 * (size_t)particles + 3 * (size_t)bonds + 6 * (size_t)angles + 12 * (size_t)dihedrals) +
           sizeof(int) * ((size_t)particles + 1) +
           sizeof(uint32_t) * (2 * (size_t)bonds + 3 * (size_t)angles + 4 * (size_t)dihedrals);
}

// F32: bonds and angles evaluated in FP32 (slots and the gather in FP32),
// kicks, drifts and positions in FP64 as always.  DIH: the block also holds
// dihedral terms, evaluated in FP64 (dihedral_core, the arithmetic of the
// per-term dihedral kernel) into slots of the gather's precision.
// Occupancy hints: a 26 KB block fits six times in a CU's LDS, i.e. six waves
// per SIMD; the FP32 one-particle-per-thread form fits 80 VGPRs without
// spills (float2 term parameters, 32-bit angle indices), so registers allow
// the same six.  The FP64 dihedral term with its dipole outputs needs about
// 200 registers: two waves per SIMD.
template <int PPT, bool F32, bool DIH>
__global__ void __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(DIH ? (PPT == 1 ? 2 : 1) : (PPT == 1 ? (F32 ? 6 : 4) : 3))))
fused_inner_kernel(FusedArgs A) {
    using R = typename std::conditional<F32, float, double>::type;
    extern __shared__ double smem_d[];
    const int blk = blockIdx.x;
    const int p0 = A.blk_start[blk];
    const int np = A.blk_start[blk + 1] - p0;
    const int b0 = A.blk_bond_off[blk], nb = A.blk_bond_off[blk + 1] - b0;
    const int a0 = A.blk_ang_off[blk], na = A.blk_ang_off[blk + 1] - a0;
    const int d0 = DIH ? A.blk_dih_off[blk] : 0, nd = DIH ? A.blk_dih_off[blk + 1] - d0 : 0;
    const int e0 = A.csr_off[p0], ne = A.csr_off[p0 + np] - e0;
    // LDS, laid out for this block's own counts: positions, bond slots (fa),
    // angle slots (fa, fc), dihedral slots (4 roles), entry offsets, entries
    double *xs = smem_d;                    // 3 * np
    R *sb = (R *)(xs + 3 * np);             // 3 * nb
    R *sa = sb + 3 * nb;                    // 6 * na
    R *sd = sa + 6 * na;                    // 12 * nd
    int *coff = (int *)(sd + 12 * nd);      // np + 1
    uint32_t *ent = (uint32_t *)(coff + np + 1);  // ne = 2 nb + 3 na + 4 nd
    for (int j = threadIdx.x; j <= np; j += blockDim.x) coff[j] = A.csr_off[p0 + j] - e0;
    for (int j = threadIdx.x; j < ne; j += blockDim.x) ent[j] = A.csr_ent[e0 + j];
    Box box;
    for (int d = 0; d < 3; ++d) {
        box.l[d] = A.box[d];
        box.inv[d] = 1.0 / A.box[d];
    }

    // Own particles: positions live in LDS only, velocities and the fast
    // force in registers.  The outer (slow) half kicks of this step, and the
    // pending one of the previous step, come first, in the order of the
    // unfused path.
    double v[PPT][3];
    float ff[PPT][3];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int j = threadIdx.x + k * kBlock;
        if (j < np) {
            const int64_t g = 3 * (int64_t)(p0 + j);
            for (int d = 0; d < 3; ++d) {
                const double fs = (double)A.f_slow[g + d];
                double vi = A.v[g + d];
                if (A.pending) vi = vi + A.h_out * fs;
                v[k][d] = vi + A.h_out * fs;
                ff[k][d] = A.f_fast[g + d];
                xs[3 * j + d] = A.x[g + d];
            }
        }
    }
    // own terms in registers, their parameters in the terms' precision (the
    // FP64 energies at the end read them again)
    using P2 = typename std::conditional<F32, float2, double2>::type;
    uint32_t bab[PPT];
    P2 bpar[PPT];
    uint32_t aabc[PPT];
    P2 apar[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int tb = threadIdx.x + k * kBlock;
        if (tb < nb) {
            bab[k] = A.bond_ab[b0 + tb];
            const double2 q = A.bond_par[b0 + tb];
            bpar[k].x = q.x;
            bpar[k].y = q.y;
        }
        if (tb < na) {
            aabc[k] = A.ang_abc[a0 + tb];
            const double2 q = A.ang_par[a0 + tb];
            apar[k].x = q.x;
            apar[k].y = q.y;
        }
    }

    double edih[2] = {0.0, 0.0};  // dihedral energy and clamps at the last sub-step
    for (int it = 0; it < A.n_inner; ++it) {
        // kicks + drift + wrap of own particles (each thread rewrites only
        // its own positions, so no barrier is needed before)
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int j = threadIdx.x + k * kBlock;
            if (j >= np) continue;
            for (int d = 0; d < 3; ++d) {
                const double vi = v[k][d] + A.h_in * (double)ff[k][d];
                v[k][d] = vi;
                xs[3 * j + d] = np_mod(xs[3 * j + d] + A.dt * vi, A.box[d]);
            }
        }
        __syncthreads();
        // every term once into its slot
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int tb = threadIdx.x + k * kBlock;
            if (tb < nb) {
                const int ia = bab[k] & 0xffff, ib = bab[k] >> 16;
                const V3 rab = mic(v3(xs[3 * ib] - xs[3 * ia], xs[3 * ib + 1] - xs[3 * ia + 1],
                                      xs[3 * ib + 2] - xs[3 * ia + 2]),
                                   box);
                if constexpr (F32) {
                    const F3 fa = bond_term_f(to_f3(rab), bpar[k].x, bpar[k].y);
                    sb[3 * tb] = fa.x;
                    sb[3 * tb + 1] = fa.y;
                    sb[3 * tb + 2] = fa.z;
                } else {
                    double en;
                    const V3 fa = bond_term(rab, bpar[k].x, bpar[k].y, en);
                    sb[3 * tb] = fa.x;
                    sb[3 * tb + 1] = fa.y;
                    sb[3 * tb + 2] = fa.z;
                }
            }
            if (tb < na) {
                const int ia = (int)(aabc[k] & 0x3ff), ib = (int)((aabc[k] >> 10) & 0x3ff),
                          ic = (int)(aabc[k] >> 20);
                const V3 rb = v3(xs[3 * ib], xs[3 * ib + 1], xs[3 * ib + 2]);
                const V3 ra = mic(v3(xs[3 * ia], xs[3 * ia + 1], xs[3 * ia + 2]) - rb, box);
                const V3 rc = mic(v3(xs[3 * ic], xs[3 * ic + 1], xs[3 * ic + 2]) - rb, box);
                if constexpr (F32) {
                    F3 fa, fc;
                    angle_term_f(ra, rc, apar[k].x, apar[k].y, fa, fc);
                    sa[6 * tb] = fa.x;
                    sa[6 * tb + 1] = fa.y;
                    sa[6 * tb + 2] = fa.z;
                    sa[6 * tb + 3] = fc.x;
                    sa[6 * tb + 4] = fc.y;
                    sa[6 * tb + 5] = fc.z;
                } else {
                    V3 fa, fc;
                    double en;
                    angle_term(ra, rc, apar[k].x, apar[k].y, fa, fc, en);
                    sa[6 * tb] = fa.x;
                    sa[6 * tb + 1] = fa.y;
                    sa[6 * tb + 2] = fa.z;
                    sa[6 * tb + 3] = fc.x;
                    sa[6 * tb + 4] = fc.y;
                    sa[6 * tb + 5] = fc.z;
                }
            }
            if constexpr (DIH) {
                if (tb < nd) {
                    // the term's constants are re-read every sub-step: an empty asm on
                    // the pointer keeps the compiler from hoisting the 27 FP64 loads out
                    // of the sub-step loop into registers held by the whole kernel
                    const double *ser = A.dih_ser + (int64_t)kDihSeries * d0 + tb;
                    asm volatile("" : "+v"(ser));
                    const uint2 q = A.dih_idx[d0 + tb];
                    const int ia = (int)(q.x & 0x3ff), ib = (int)((q.x >> 10) & 0x3ff), ic = (int)(q.x >> 20),
                              id = (int)(q.y & 0x3ff), dt = (int)((q.y >> 10) & 3), last = (int)((q.y >> 12) & 1);
                    const bool final_step = it == A.n_inner - 1;
                    double evd[2] = {0.0, 0.0};
                    V3 out[4];
                    // forces and energy only (the dipole outputs' FP64 transfer matrices
                    // would double this kernel's registers); the cosine series from
                    // their constants by the harmonic recurrence
                    dihedral_core_series(v3(xs[3 * ia], xs[3 * ia + 1], xs[3 * ia + 2]),
                                         v3(xs[3 * ib], xs[3 * ib + 1], xs[3 * ib + 2]),
                                         v3(xs[3 * ic], xs[3 * ic + 1], xs[3 * ic + 2]),
                                         v3(xs[3 * id], xs[3 * id + 1], xs[3 * id + 2]),
                                         ser, nd,
                                         A.dih_coeff + 30 * (int64_t)(d0 + tb), dt, last, box, evd, out);
                    if (final_step) {
                        edih[0] += evd[0];
                        edih[1] += evd[1];
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        sd[12 * tb + 3 * r] = (R)out[r].x;
                        sd[12 * tb + 3 * r + 1] = (R)out[r].y;
                        sd[12 * tb + 3 * r + 2] = (R)out[r].z;
                    }
                }
            }
        }
        __syncthreads();
        // gather own particles' contributions in term order
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int j = threadIdx.x + k * kBlock;
            if (j >= np) continue;
            R acc[3] = {R(0), R(0), R(0)};
            for (int q = coff[j]; q < coff[j + 1]; ++q) {
                const uint32_t e = ent[q];
                const int role = e & 3, kind = (e >> 2) & 3, t = (int)(e >> 4);
                if (kind == 0) {
                    const R s = role == 0 ? R(-1) : R(1);
                    for (int d = 0; d < 3; ++d) acc[d] += s * sb[3 * t + d];
                } else if (!DIH || kind == 1) {
                    for (int d = 0; d < 3; ++d) {
                        const R fa = sa[6 * t + d], fc = sa[6 * t + 3 + d];
                        acc[d] += role == 0 ? -fa : (role == 2 ? -fc : fa + fc);
                    }
                } else {
                    for (int d = 0; d < 3; ++d) acc[d] += sd[12 * t + 3 * role + d];
                }
            }
            // the sub-step's second half kick (the next sub-step's first
            // adds the same term again: two FP64 additions, as before)
            for (int d = 0; d < 3; ++d) {
                ff[k][d] = (float)acc[d];
                v[k][d] = v[k][d] + A.h_in * (double)ff[k][d];
            }
        }
        // slots are rewritten only after the next sub-step's barrier
    }
    // write back (the last sub-step's second half kick is in v)
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int j = threadIdx.x + k * kBlock;
        if (j >= np) continue;
        const int64_t g = 3 * (int64_t)(p0 + j);
        for (int d = 0; d < 3; ++d) {
            const double xi = xs[3 * j + d];
            A.v[g + d] = v[k][d];
            A.x[g + d] = xi;
            A.x32[g + d] = (float)xi;
            A.f_fast[g + d] = ff[k][d];
        }
    }
    if (A.key.keys) {  // the field path's row keys of the float32 positions
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int j = threadIdx.x + k * kBlock;
            int row = -1;
            if (j < np) {
                const CicAxis cx = cic_axis((double)(float)xs[3 * j], A.key.inv_h[0], A.key.mx);
                const CicAxis cy = cic_axis((double)(float)xs[3 * j + 1], A.key.inv_h[1], A.key.my);
                row = cx.i0 * A.key.my + cy.i0;
            }
            const int rank = rank_in_bin(row, A.key.counts);
            if (j < np) A.key.keys[p0 + j] = make_int2(row, rank);
        }
    }
    if (A.partials) {
        // energies and virials at the positions of the last force evaluation
        // (unchanged since: the final kick moves velocities only)
        double ev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
            const int tb = threadIdx.x + k * kBlock;
            if (tb < nb) {
                const int ia = bab[k] & 0xffff, ib = bab[k] >> 16;
                const V3 rab = mic(v3(xs[3 * ib] - xs[3 * ia], xs[3 * ib + 1] - xs[3 * ia + 1],
                                      xs[3 * ib + 2] - xs[3 * ia + 2]),
                                   box);
                double en;
                const double2 q = A.bond_par[b0 + tb];
                const V3 fa = bond_term(rab, q.x, q.y, en);
                ev[0] += en;
                ev[1] += fa.x * rab.x;
                ev[2] += fa.y * rab.y;
                ev[3] += fa.z * rab.z;
            }
            if (tb < na) {
                const int ia = (int)(aabc[k] & 0x3ff), ib = (int)((aabc[k] >> 10) & 0x3ff),
                          ic = (int)(aabc[k] >> 20);
                const V3 rb = v3(xs[3 * ib], xs[3 * ib + 1], xs[3 * ib + 2]);
                const V3 ra = mic(v3(xs[3 * ia], xs[3 * ia + 1], xs[3 * ia + 2]) - rb, box);
                const V3 rc = mic(v3(xs[3 * ic], xs[3 * ic + 1], xs[3 * ic + 2]) - rb, box);
                V3 fa, fc;
                double en;
                const double2 q = A.ang_par[a0 + tb];
                angle_term(ra, rc, q.x, q.y, fa, fc, en);
                ev[4] += en;
                ev[5] -= fa.x * ra.x + fc.x * rc.x;
                ev[6] -= fa.y * ra.y + fc.y * rc.y;
                ev[7] -= fa.z * ra.z + fc.z * rc.z;
            }
        }
        if constexpr (DIH) {
            double ev10[10];
            for (int j = 0; j < 8; ++j) ev10[j] = ev[j];
            ev10[8] = edih[0];
            ev10[9] = edih[1];
            block_reduce_store<10>(ev10, A.partials + 10 * blk);
        } else {
            block_reduce_store<8>(ev, A.partials + 8 * blk);
        }
    }
}

}  // namespace hymd

using namespace hymd;

struct hymd_fused {
    int64_t n = 0;
    int nblocks = 0, ppt = 1;
    int max_bonds = 0, max_angles = 0, max_entries = 0, max_particles = 0;
    int64_t single_tail = 0;  // rows [single_tail, n) are molecules of one particle (no terms)
    size_t lds = 0;
    bool f32 = false;         // bond / angle terms in FP32
    int64_t m4 = 0;           // dihedral terms (0: the kernel without them)
    DevBuf blk_start, blk_bond_off, bond_ab, bond_par, blk_ang_off, ang_abc, ang_par, csr_off, csr_ent;
    DevBuf blk_dih_off, dih_idx, dih_coeff, dih_ser;
    DevBuf partials, result;
    HostBuf host;
};

namespace {

struct UF {
    std::vector<int> p;
    explicit UF(int64_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
    int find(int a) {
        while (p[a] != a) a = p[a] = p[p[a]];
        return a;
    }
    void join(int a, int b) {
        a = find(a);
        b = find(b);
        if (a != b) p[std::max(a, b)] = std::min(a, b);
    }
};

template <typename T>
int upload(DevBuf &buf, const std::vector<T> &h) {
    HYMD_TRY(buf.reserve(sizeof(T) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HYMD_HIP(hipMemcpy(buf.ptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return HYMD_OK;
}

template <typename T>
int download(std::vector<T> &h, const void *d, int64_t n) {
    h.resize(n);
    if (n > 0) HYMD_HIP(hipMemcpy(h.data(), d, sizeof(T) * n, hipMemcpyDeviceToHost));
    return HYMD_OK;
}

}  // namespace

// Builds the blocks from device term arrays.  Returns HYMD_ERR_UNSUPPORTED
// (with a message) when a molecule is not a contiguous index range or does
// not fit one workgroup; the caller then keeps the unfused kernels.
namespace hymd {
int64_t fused_single_tail(const hymd_fused *f) { return f ? f->single_tail : 0; }
}  // namespace hymd

extern "C" int hymd_fused_create_dihedrals(hymd_fused **out, int64_t n, int64_t m2, const int32_t *a2,
                                           const int32_t *b2, const double *r0, const double *k2, int64_t m3,
                                           const int32_t *a3, const int32_t *b3, const int32_t *c3, const double *t0,
                                           const double *k3, int64_t m4, const int32_t *a4, const int32_t *b4,
                                           const int32_t *c4, const int32_t *d4, const float *coeff,
                                           const int32_t *dtype, const int32_t *last4) {
    HYMD_REQUIRE(out && n >= 0 && m2 >= 0 && m3 >= 0 && m4 >= 0 && n < (int64_t(1) << 31) &&
                     m4 < (int64_t(1) << 28),
                 "bad argument");
    HYMD_REQUIRE(m4 == 0 || (a4 && b4 && c4 && d4 && coeff && dtype && last4), "null dihedral array");
    std::vector<int32_t> A2, B2, A3, B3, C3, A4, B4, C4, D4, DT, LA;
    std::vector<double> R0, K2, T0, K3;
    std::vector<float> CF;
    HYMD_TRY(download(A2, a2, m2));
    HYMD_TRY(download(B2, b2, m2));
    HYMD_TRY(download(R0, r0, m2));
    HYMD_TRY(download(K2, k2, m2));
    HYMD_TRY(download(A3, a3, m3));
    HYMD_TRY(download(B3, b3, m3));
    HYMD_TRY(download(C3, c3, m3));
    HYMD_TRY(download(T0, t0, m3));
    HYMD_TRY(download(K3, k3, m3));
    HYMD_TRY(download(A4, a4, m4));
    HYMD_TRY(download(B4, b4, m4));
    HYMD_TRY(download(C4, c4, m4));
    HYMD_TRY(download(D4, d4, m4));
    HYMD_TRY(download(DT, dtype, m4));
    HYMD_TRY(download(LA, last4, m4));
    HYMD_TRY(download(CF, coeff, 30 * m4));
    auto bad = [n](int32_t i) { return i < 0 || (int64_t)i >= n; };
    UF uf(n);
    for (int64_t t = 0; t < m2; ++t) {
        HYMD_REQUIRE(!bad(A2[t]) && !bad(B2[t]), "bond %lld has an index outside [0, %lld)", (long long)t, (long long)n);
        uf.join(A2[t], B2[t]);
    }
    for (int64_t t = 0; t < m3; ++t) {
        HYMD_REQUIRE(!bad(A3[t]) && !bad(B3[t]) && !bad(C3[t]), "angle %lld has an index outside [0, %lld)",
                     (long long)t, (long long)n);
        uf.join(A3[t], B3[t]);
        uf.join(B3[t], C3[t]);
    }
    for (int64_t t = 0; t < m4; ++t) {
        HYMD_REQUIRE(!bad(A4[t]) && !bad(B4[t]) && !bad(C4[t]) && !bad(D4[t]),
                     "dihedral %lld has an index outside [0, %lld)", (long long)t, (long long)n);
        HYMD_REQUIRE(DT[t] >= 0 && DT[t] <= 2 && (LA[t] == 0 || LA[t] == 1),
                     "dihedral %lld has type %d / last %d (types 0-2, last 0 or 1)", (long long)t, DT[t], LA[t]);
        uf.join(A4[t], B4[t]);
        uf.join(B4[t], C4[t]);
        uf.join(C4[t], D4[t]);
    }
    // components must be contiguous: root = smallest member, last member = root + size - 1
    std::vector<int> size(n, 0), last(n, -1);
    for (int64_t i = 0; i < n; ++i) {
        const int r = uf.find((int)i);
        size[r]++;
        last[r] = (int)i;
    }
    int max_mol = 0;
    for (int64_t i = 0; i < n; ++i)
        if (uf.find((int)i) == (int)i) max_mol = std::max(max_mol, size[i]);
    const int ppt = max_mol <= kBlock ? 1 : 2;
    const int pmax = kBlock * ppt;
    const size_t kFusedLds = m4 > 0 ? (ppt == 1 ? kFusedLdsDih1 : kFusedLdsDih2) : (ppt == 1 ? kFusedLds1 : kFusedLds2);
    // terms per molecule (indexed by its root)
    std::vector<int> mol_b(n, 0), mol_a(n, 0), mol_d(n, 0);
    for (int64_t t = 0; t < m2; ++t) mol_b[uf.find(A2[t])]++;
    for (int64_t t = 0; t < m3; ++t) mol_a[uf.find(A3[t])]++;
    for (int64_t t = 0; t < m4; ++t) mol_d[uf.find(A4[t])]++;
    // Pack whole molecules into blocks of at most pmax particles, kBlock *
    // ppt terms of each kind, and kFusedLds bytes of LDS.
    std::vector<int> blk_start{0};
    int cur_p = 0, cur_b = 0, cur_a = 0, cur_d = 0;
    const int tmax = kBlock * ppt;
    for (int64_t i = 0; i < n;) {
        const int r = uf.find((int)i);
        if (r != (int)i || last[r] - r + 1 != size[r]) {
            set_error("molecule containing particle %lld is not a contiguous index range", (long long)i);
            return HYMD_ERR_UNSUPPORTED;
        }
        if (size[r] > pmax || mol_b[r] > tmax || mol_a[r] > tmax || mol_d[r] > tmax ||
            fused_lds_bytes(size[r], mol_b[r], mol_a[r], mol_d[r]) > kFusedLds) {
            set_error("a molecule of %d beads, %d bonds, %d angles and %d dihedrals exceeds one fused block",
                      size[r], mol_b[r], mol_a[r], mol_d[r]);
            return HYMD_ERR_UNSUPPORTED;
        }
        if (cur_p + size[r] > pmax || cur_b + mol_b[r] > tmax || cur_a + mol_a[r] > tmax ||
            cur_d + mol_d[r] > tmax ||
            fused_lds_bytes(cur_p + size[r], cur_b + mol_b[r], cur_a + mol_a[r], cur_d + mol_d[r]) > kFusedLds) {
            blk_start.push_back((int)i);
            cur_p = cur_b = cur_a = cur_d = 0;
        }
        cur_p += size[r];
        cur_b += mol_b[r];
        cur_a += mol_a[r];
        cur_d += mol_d[r];
        i += size[r];
    }
    if (blk_start.back() != (int)n) blk_start.push_back((int)n);
    const int nblk = (int)blk_start.size() - 1;
    std::vector<int> blk_of(n);
    for (int b = 0; b < nblk; ++b)
        for (int i = blk_start[b]; i < blk_start[b + 1]; ++i) blk_of[i] = b;
    // terms per block, in global term order
    std::vector<int> bcount(nblk + 1, 0), acount(nblk + 1, 0), dcount(nblk + 1, 0);
    for (int64_t t = 0; t < m2; ++t) bcount[blk_of[A2[t]] + 1]++;
    for (int64_t t = 0; t < m3; ++t) acount[blk_of[A3[t]] + 1]++;
    for (int64_t t = 0; t < m4; ++t) dcount[blk_of[A4[t]] + 1]++;
    std::partial_sum(bcount.begin(), bcount.end(), bcount.begin());
    std::partial_sum(acount.begin(), acount.end(), acount.begin());
    std::partial_sum(dcount.begin(), dcount.end(), dcount.begin());
    std::vector<uint32_t> bab(m2);
    std::vector<double2> bpar(m2);
    std::vector<uint32_t> aabc(m3);
    std::vector<double2> apar(m3);
    std::vector<int> bfill(bcount.begin(), bcount.end() - 1), afill(acount.begin(), acount.end() - 1);
    std::vector<int> bond_local(m2), ang_local(m3), dih_local(m4);
    std::vector<uint2> didx(m4);
    std::vector<float> dcf(30 * m4);
    std::vector<double> dser((size_t)kDihSeries * m4);
    std::vector<int> dfill(dcount.begin(), dcount.end() - 1);
    for (int64_t t = 0; t < m2; ++t) {
        const int b = blk_of[A2[t]], s = blk_start[b], slot = bfill[b]++;
        bab[slot] = (uint32_t)(A2[t] - s) | ((uint32_t)(B2[t] - s) << 16);
        bpar[slot] = make_double2(R0[t], K2[t]);
        bond_local[t] = slot - bcount[b];
    }
    for (int64_t t = 0; t < m3; ++t) {
        const int b = blk_of[A3[t]], s = blk_start[b], slot = afill[b]++;
        aabc[slot] = (uint32_t)(A3[t] - s) | ((uint32_t)(B3[t] - s) << 10) | ((uint32_t)(C3[t] - s) << 20);
        apar[slot] = make_double2(T0[t], K3[t]);
        ang_local[t] = slot - acount[b];
    }
    for (int64_t t = 0; t < m4; ++t) {
        const int b = blk_of[A4[t]], s = blk_start[b], slot = dfill[b]++;
        didx[slot] = make_uint2((uint32_t)(A4[t] - s) | ((uint32_t)(B4[t] - s) << 10) | ((uint32_t)(C4[t] - s) << 20),
                                (uint32_t)(D4[t] - s) | ((uint32_t)DT[t] << 10) | ((uint32_t)LA[t] << 12));
        std::copy(CF.begin() + 30 * t, CF.begin() + 30 * (t + 1), dcf.begin() + 30 * (int64_t)slot);
        double k27[kDihSeries];
        dihedral_series_constants(&CF[30 * t], k27);
        const int d0 = dcount[b], nd = dcount[b + 1] - dcount[b], j = slot - d0;
        for (int k = 0; k < kDihSeries; ++k) dser[(size_t)kDihSeries * d0 + (size_t)k * nd + j] = k27[k];
        dih_local[t] = slot - dcount[b];
    }
    int max_b = 0, max_a = 0, max_p = 0, max_d = 0;
    for (int b = 0; b < nblk; ++b) {
        max_b = std::max(max_b, bcount[b + 1] - bcount[b]);
        max_a = std::max(max_a, acount[b + 1] - acount[b]);
        max_d = std::max(max_d, dcount[b + 1] - dcount[b]);
        max_p = std::max(max_p, blk_start[b + 1] - blk_start[b]);
    }
    if (max_b > tmax || max_a > tmax || max_d > tmax) {
        set_error("too many bonded terms per block for the fused kernel");
        return HYMD_ERR_UNSUPPORTED;
    }
    // per-particle entries: bonds in term order, then angles, then dihedrals
    std::vector<int> off(n + 1, 0);
    for (int64_t t = 0; t < m2; ++t) {
        off[A2[t] + 1]++;
        off[B2[t] + 1]++;
    }
    for (int64_t t = 0; t < m3; ++t) {
        off[A3[t] + 1]++;
        off[B3[t] + 1]++;
        off[C3[t] + 1]++;
    }
    for (int64_t t = 0; t < m4; ++t) {
        off[A4[t] + 1]++;
        off[B4[t] + 1]++;
        off[C4[t] + 1]++;
        off[D4[t] + 1]++;
    }
    std::partial_sum(off.begin(), off.end(), off.begin());
    std::vector<uint32_t> ent(off[n]);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int64_t t = 0; t < m2; ++t) {
        const uint32_t base = (uint32_t)bond_local[t] << 4;
        ent[fill[A2[t]]++] = base | 0u;
        ent[fill[B2[t]]++] = base | 1u;
    }
    for (int64_t t = 0; t < m3; ++t) {
        const uint32_t base = ((uint32_t)ang_local[t] << 4) | 4u;
        ent[fill[A3[t]]++] = base | 0u;
        ent[fill[B3[t]]++] = base | 1u;
        ent[fill[C3[t]]++] = base | 2u;
    }
    for (int64_t t = 0; t < m4; ++t) {
        const uint32_t base = ((uint32_t)dih_local[t] << 4) | 8u;
        ent[fill[A4[t]]++] = base | 0u;
        ent[fill[B4[t]]++] = base | 1u;
        ent[fill[C4[t]]++] = base | 2u;
        ent[fill[D4[t]]++] = base | 3u;
    }
    int max_e = 0;
    for (int b = 0; b < nblk; ++b) max_e = std::max(max_e, off[blk_start[b + 1]] - off[blk_start[b]]);
    hymd_fused *f = new hymd_fused();
    f->n = n;
    f->single_tail = n;
    while (f->single_tail > 0 && size[uf.find((int)(f->single_tail - 1))] == 1) --f->single_tail;
    f->nblocks = nblk;
    f->ppt = ppt;
    f->max_bonds = std::max(max_b, 1);
    f->max_angles = std::max(max_a, 1);
    f->max_entries = std::max(max_e, 1);
    f->max_particles = std::max(max_p, 1);
    f->m4 = m4;
    f->lds = 0;
    for (int b = 0; b < nblk; ++b)
        f->lds = std::max(f->lds, fused_lds_bytes(blk_start[b + 1] - blk_start[b], bcount[b + 1] - bcount[b],
                                                  acount[b + 1] - acount[b], dcount[b + 1] - dcount[b]));
    int st = HYMD_OK;
    if (f->lds > kFusedLds) {
        set_error("fused block needs %zu bytes of LDS", f->lds);
        st = HYMD_ERR_UNSUPPORTED;
    }
    if (st == HYMD_OK) st = upload(f->blk_start, blk_start);
    if (st == HYMD_OK) st = upload(f->blk_bond_off, bcount);
    if (st == HYMD_OK) st = upload(f->bond_ab, bab);
    if (st == HYMD_OK) st = upload(f->bond_par, bpar);
    if (st == HYMD_OK) st = upload(f->blk_ang_off, acount);
    if (st == HYMD_OK) st = upload(f->ang_abc, aabc);
    if (st == HYMD_OK) st = upload(f->ang_par, apar);
    if (st == HYMD_OK) st = upload(f->csr_off, off);
    if (st == HYMD_OK) st = upload(f->csr_ent, ent);
    if (st == HYMD_OK && m4 > 0) st = upload(f->blk_dih_off, dcount);
    if (st == HYMD_OK && m4 > 0) st = upload(f->dih_idx, didx);
    if (st == HYMD_OK && m4 > 0) st = upload(f->dih_coeff, dcf);
    if (st == HYMD_OK && m4 > 0) st = upload(f->dih_ser, dser);
    if (st == HYMD_OK) st = f->partials.reserve(sizeof(double) * 10 * std::max(nblk, 1));
    if (st == HYMD_OK) st = f->result.reserve(sizeof(double) * 10);
    if (st == HYMD_OK) st = f->host.reserve(sizeof(double) * 10);
    if (st != HYMD_OK) {
        delete f;
        return st;
    }
    *out = f;
    return HYMD_OK;
}

extern "C" int hymd_fused_create(hymd_fused **out, int64_t n, int64_t m2, const int32_t *a2, const int32_t *b2,
                                 const double *r0, const double *k2, int64_t m3, const int32_t *a3,
                                 const int32_t *b3, const int32_t *c3, const double *t0, const double *k3) {
    return hymd_fused_create_dihedrals(out, n, m2, a2, b2, r0, k2, m3, a3, b3, c3, t0, k3, 0, nullptr, nullptr,
                                       nullptr, nullptr, nullptr, nullptr, nullptr);
}

extern "C" int hymd_fused_destroy(hymd_fused *f) {
    delete f;
    return HYMD_OK;
}

namespace hymd {
// Dihedral blocks of molecules above 256 beads exceed the default 64 KB of
// dynamic LDS: opt in once per kernel and process.
static int fused_lds_opt_in(const hymd_fused *f) {
    static bool done[2][2] = {{false, false}, {false, false}};
    bool &d = done[f->ppt - 1][f->f32 ? 1 : 0];
    if (d) return HYMD_OK;
    const void *k = f->f32 ? (f->ppt == 1 ? reinterpret_cast<const void *>(&fused_inner_kernel<1, true, true>)
                                          : reinterpret_cast<const void *>(&fused_inner_kernel<2, true, true>))
                           : (f->ppt == 1 ? reinterpret_cast<const void *>(&fused_inner_kernel<1, false, true>)
                                          : reinterpret_cast<const void *>(&fused_inner_kernel<2, false, true>));
    HYMD_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFusedLdsDih2));
    d = true;
    return HYMD_OK;
}

int fused_inner_impl(hymd_fused *f, double *x, double *v, float *x32, float *f_fast, const float *f_slow, double h_in,
                     double h_out, double dt, int32_t n_inner, int32_t pending, const double box[3], double *energies,
                     double *virials, void *stream, const RowKeyTarget *key, const FusedDihOut *dih) {
    HYMD_REQUIRE(f && box && n_inner >= 1 && (f->n == 0 || (x && v && x32 && f_fast && f_slow)), "bad argument");
    hipStream_t s = (hipStream_t)stream;
    FusedArgs A;
    A.x = x;
    A.v = v;
    A.x32 = x32;
    A.f_fast = f_fast;
    A.f_slow = f_slow;
    A.blk_start = (const int *)f->blk_start.ptr;
    A.blk_bond_off = (const int *)f->blk_bond_off.ptr;
    A.bond_ab = (const uint32_t *)f->bond_ab.ptr;
    A.bond_par = (const double2 *)f->bond_par.ptr;
    A.blk_ang_off = (const int *)f->blk_ang_off.ptr;
    A.ang_abc = (const uint32_t *)f->ang_abc.ptr;
    A.ang_par = (const double2 *)f->ang_par.ptr;
    A.csr_off = (const int *)f->csr_off.ptr;
    A.csr_ent = (const uint32_t *)f->csr_ent.ptr;
    A.blk_dih_off = (const int *)f->blk_dih_off.ptr;
    A.dih_idx = (const uint2 *)f->dih_idx.ptr;
    A.dih_coeff = (const float *)f->dih_coeff.ptr;
    A.dih_ser = (const double *)f->dih_ser.ptr;
    A.h_in = h_in;
    A.h_out = h_out;
    A.dt = dt;
    for (int d = 0; d < 3; ++d) A.box[d] = box[d];
    A.n_inner = n_inner;
    A.pending = pending ? 1 : 0;
    const bool red = energies || virials || (dih && dih->energy);
    const int nv = f->m4 > 0 ? 10 : 8;
    A.partials = red ? (double *)f->partials.ptr : nullptr;
    if (key) A.key = *key;
    else A.key.keys = nullptr;
    if (f->nblocks > 0) {
        if (f->m4 > 0) {
            if (f->lds > 64 * 1024) HYMD_TRY(fused_lds_opt_in(f));
            if (f->f32) {
                if (f->ppt == 1) fused_inner_kernel<1, true, true><<<f->nblocks, kBlock, f->lds, s>>>(A);
                else fused_inner_kernel<2, true, true><<<f->nblocks, kBlock, f->lds, s>>>(A);
            } else {
                if (f->ppt == 1) fused_inner_kernel<1, false, true><<<f->nblocks, kBlock, f->lds, s>>>(A);
                else fused_inner_kernel<2, false, true><<<f->nblocks, kBlock, f->lds, s>>>(A);
            }
        } else if (f->f32) {
            if (f->ppt == 1) fused_inner_kernel<1, true, false><<<f->nblocks, kBlock, f->lds, s>>>(A);
            else fused_inner_kernel<2, true, false><<<f->nblocks, kBlock, f->lds, s>>>(A);
        } else {
            if (f->ppt == 1) fused_inner_kernel<1, false, false><<<f->nblocks, kBlock, f->lds, s>>>(A);
            else fused_inner_kernel<2, false, false><<<f->nblocks, kBlock, f->lds, s>>>(A);
        }
    }
    HYMD_HIP(hipGetLastError());
    if (red) {
        HYMD_TRY(reduce_partials((const double *)f->partials.ptr, f->nblocks, nv, (double *)f->result.ptr, s));
        HYMD_HIP(hipMemcpyAsync(f->host.ptr, f->result.ptr, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
        HYMD_HIP(hipStreamSynchronize(s));
        const double *h = (const double *)f->host.ptr;
        if (energies) {
            energies[0] = h[0];
            energies[1] = h[4];
        }
        if (virials)
            for (int d = 0; d < 3; ++d) {
                virials[d] = h[1 + d];
                virials[3 + d] = h[5 + d];
            }
        if (dih && dih->energy) {
            dih->energy[0] = f->m4 > 0 ? h[8] : 0.0;
            dih->energy[1] = f->m4 > 0 ? h[9] : 0.0;
        }
    }
    return HYMD_OK;
}
}  // namespace hymd

extern "C" int hymd_fused_set_precision(hymd_fused *f, int32_t precision) {
    HYMD_REQUIRE(f, "null handle");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    f->f32 = precision == HYMD_F32;
    return HYMD_OK;
}

extern "C" int hymd_fused_inner(hymd_fused *f, double *x, double *v, float *x32, float *f_fast, const float *f_slow,
                                double h_in, double h_out, double dt, int32_t n_inner, int32_t pending,
                                const double box[3], double *energies, double *virials, void *stream) {
    return fused_inner_impl(f, x, v, x32, f_fast, f_slow, h_in, h_out, dt, n_inner, pending, box, energies, virials,
                            stream, nullptr, nullptr);
}

extern "C" int hymd_fused_inner_dihedrals(hymd_fused *f, double *x, double *v, float *x32, float *f_fast,
                                          const float *f_slow, double h_in, double h_out, double dt, int32_t n_inner,
                                          int32_t pending, const double box[3], double *energies, double *virials,
                                          void *stream) {
    double eb[2], ed[2];
    FusedDihOut out{energies ? ed : nullptr};
    HYMD_TRY(fused_inner_impl(f, x, v, x32, f_fast, f_slow, h_in, h_out, dt, n_inner, pending, box,
                              energies ? eb : nullptr, virials, stream, nullptr, &out));
    if (energies) {
        energies[0] = eb[0];
        energies[1] = eb[1];
        energies[2] = ed[0];
        energies[3] = ed[1];
    }
    return HYMD_OK;
}
