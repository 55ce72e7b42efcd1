// Bonded kernels (cbf / caf / cdf and their _d twins), dipole force
// redistribution, and the persistent topology used by the rRESPA loop.
//
// Per-particle term lists (CSR) are built on the device: count members,
// exclusive scan, fill with atomics, then sort every short list by
// (term, role) so each particle sums its contributions in the reference's
// term order.  One thread per particle then recomputes its terms in FP64.
#include <cmath>
#include <memory>
#include <mutex>

#include "bonded.h"

namespace hymd {

// ---------------------------------------------------------------------------
// CSR build
// ---------------------------------------------------------------------------
struct TermIdx {
    const int32_t *m[4];  // member arrays (roles 0..nroles-1)
    int nroles;
};

__global__ void csr_count_kernel(TermIdx ti, int64_t m, int64_t n, int *__restrict__ counts, int *__restrict__ err) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride)
        for (int r = 0; r < ti.nroles; ++r) {
            const int32_t p = ti.m[r][t];
            if (p < 0 || (int64_t)p >= n) {
                atomicOr(err, 1);
                continue;
            }
            atomicAdd(counts + p, 1);
        }
}

__global__ void csr_fill_kernel(TermIdx ti, int64_t m, int64_t n, int *__restrict__ cursor,
                                uint32_t *__restrict__ entries) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride)
        for (int r = 0; r < ti.nroles; ++r) {
            const int32_t p = ti.m[r][t];
            if (p < 0 || (int64_t)p >= n) continue;
            const int pos = atomicAdd(cursor + p, 1);
            entries[pos] = ((uint32_t)t << 2) | (uint32_t)r;
        }
}

__global__ void csr_sort_kernel(const int *__restrict__ off, int64_t n, uint32_t *__restrict__ entries) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const int b = off[p], e = off[p + 1];
        for (int i = b + 1; i < e; ++i) {
            const uint32_t key = entries[i];
            int j = i - 1;
            while (j >= b && entries[j] > key) {
                entries[j + 1] = entries[j];
                --j;
            }
            entries[j + 1] = key;
        }
    }
}

struct Csr {
    int64_t n = 0, m = 0;
    DevBuf counts;   // n ints (reused as cursor)
    DevBuf offsets;  // n+1 ints
    DevBuf entries;  // nroles*m uint32
    DevBuf tiles;    // scan tile sums + grand total + error flag
    const int *off() const { return (const int *)offsets.ptr; }
    const uint32_t *ent() const { return (const uint32_t *)entries.ptr; }
};

// Builds the lists into caller-provided storage: counts (n ints, left holding
// the offsets), off (n + 1 ints), entries (nroles * m words), tiles
// (scan_tiles(n) + 2 ints, the last one the index error flag).  Clears with a
// fill kernel, checks the indices with one synchronising read-back.
static int build_csr_into(int *counts, int *off, uint32_t *entries, int *tiles, const TermIdx &ti, int64_t m,
                          int64_t n, hipStream_t s) {
    int *err = tiles + scan_tiles(n) + 1;
    HYMD_TRY(zero_words(counts, std::max<int64_t>(n, 1), s));
    HYMD_TRY(zero_words(err, 1, s));
    if (m > 0) csr_count_kernel<<<grid_for(m), kBlock, 0, s>>>(ti, m, n, counts, err);
    HYMD_HIP(hipGetLastError());
    HYMD_TRY(exclusive_scan(counts, off, n, tiles, n > 0 ? counts : nullptr, false, s));
    if (m > 0) csr_fill_kernel<<<grid_for(m), kBlock, 0, s>>>(ti, m, n, counts, entries);
    if (n > 0) csr_sort_kernel<<<grid_for(n), kBlock, 0, s>>>(off, n, entries);
    HYMD_HIP(hipGetLastError());
    int herr = 0;
    HYMD_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    HYMD_REQUIRE(herr == 0, "a bonded-term particle index is outside [0, %lld)", (long long)n);
    return HYMD_OK;
}

static int build_csr(Csr &c, const TermIdx &ti, int64_t m, int64_t n, hipStream_t s) {
    HYMD_REQUIRE(m < (int64_t(1) << 29), "too many terms (%lld)", (long long)m);
    HYMD_REQUIRE(m * ti.nroles < (int64_t(1) << 31), "too many term members");
    c.n = n;
    c.m = m;
    HYMD_TRY(c.counts.reserve(sizeof(int) * std::max<int64_t>(n, 1)));
    HYMD_TRY(c.offsets.reserve(sizeof(int) * (n + 1)));
    HYMD_TRY(c.entries.reserve(sizeof(uint32_t) * std::max<int64_t>(m * ti.nroles, 1)));
    HYMD_TRY(c.tiles.reserve(sizeof(int) * (scan_tiles(n) + 2)));
    return build_csr_into((int *)c.counts.ptr, (int *)c.offsets.ptr, (uint32_t *)c.entries.ptr, (int *)c.tiles.ptr, ti,
                          m, n, s);
}

// ---------------------------------------------------------------------------
// force kernels (gather)
// ---------------------------------------------------------------------------
struct BondArgs {
    const int32_t *a, *b;
    const double *r0, *k;
};
struct AngleArgs {
    const int32_t *a, *b, *c;
    const double *t0, *k;
};

template <typename PosT>
__device__ __forceinline__ V3 bond_contrib(const PosT *__restrict__ r, const BondArgs &B, uint32_t e, const Box &box,
                                           double *ev) {
    const int64_t t = e >> 2;
    const int role = e & 3;
    const V3 rab = mic(load3(r, B.b[t]) - load3(r, B.a[t]), box);
    double en;
    const V3 fa = bond_term(rab, B.r0[t], B.k[t], en);
    if (role == 0) {
        ev[0] += en;
        ev[1] += fa.x * rab.x;
        ev[2] += fa.y * rab.y;
        ev[3] += fa.z * rab.z;
        return -fa;
    }
    return fa;
}

template <typename PosT>
__device__ __forceinline__ V3 angle_contrib(const PosT *__restrict__ r, const AngleArgs &A, uint32_t e, const Box &box,
                                            double *ev) {
    const int64_t t = e >> 2;
    const int role = e & 3;
    const V3 rb = load3(r, A.b[t]);
    const V3 ra = mic(load3(r, A.a[t]) - rb, box);
    const V3 rc = mic(load3(r, A.c[t]) - rb, box);
    V3 fa, fc;
    double en;
    if (!angle_term(ra, rc, A.t0[t], A.k[t], fa, fc, en)) return v3(0, 0, 0);
    if (role == 0) {
        ev[0] += en;
        ev[1] -= fa.x * ra.x + fc.x * rc.x;
        ev[2] -= fa.y * ra.y + fc.y * rc.y;
        ev[3] -= fa.z * ra.z + fc.z * rc.z;
        return -fa;
    }
    if (role == 2) return -fc;
    return fa + fc;
}

template <typename PosT, typename OutT>
__global__ void __launch_bounds__(kBlock) bond_gather_kernel(const PosT *__restrict__ r, OutT *__restrict__ f,
                                                              int64_t n, const int *__restrict__ off,
                                                              const uint32_t *__restrict__ ent, BondArgs B, Box box,
                                                              double *__restrict__ partials) {
    double ev[4] = {0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        V3 acc = v3(0, 0, 0);
        for (int j = off[p]; j < off[p + 1]; ++j) acc = acc + bond_contrib(r, B, ent[j], box, ev);
        f[3 * p] = (OutT)acc.x;
        f[3 * p + 1] = (OutT)acc.y;
        f[3 * p + 2] = (OutT)acc.z;
    }
    if (partials) block_reduce_store<4>(ev, partials + 4 * blockIdx.x);
}

template <typename PosT, typename OutT>
__global__ void __launch_bounds__(kBlock) angle_gather_kernel(const PosT *__restrict__ r, OutT *__restrict__ f,
                                                               int64_t n, const int *__restrict__ off,
                                                               const uint32_t *__restrict__ ent, AngleArgs A, Box box,
                                                               double *__restrict__ partials) {
    double ev[4] = {0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        V3 acc = v3(0, 0, 0);
        for (int j = off[p]; j < off[p + 1]; ++j) acc = acc + angle_contrib(r, A, ent[j], box, ev);
        f[3 * p] = (OutT)acc.x;
        f[3 * p + 1] = (OutT)acc.y;
        f[3 * p + 2] = (OutT)acc.z;
    }
    if (partials) block_reduce_store<4>(ev, partials + 4 * blockIdx.x);
}

// bonds + angles of the persistent topology in one pass (float64 positions,
// float32 total bonded force)
__global__ void __launch_bounds__(kBlock) topo_gather_kernel(const double *__restrict__ r, float *__restrict__ f,
                                                              int64_t n, const int *__restrict__ off2,
                                                              const uint32_t *__restrict__ ent2, BondArgs B,
                                                              const int *__restrict__ off3,
                                                              const uint32_t *__restrict__ ent3, AngleArgs A, Box box,
                                                              double *__restrict__ partials) {
    double ev[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        V3 acc = v3(0, 0, 0);
        if (off2)
            for (int j = off2[p]; j < off2[p + 1]; ++j) acc = acc + bond_contrib(r, B, ent2[j], box, ev);
        if (off3)
            for (int j = off3[p]; j < off3[p + 1]; ++j) acc = acc + angle_contrib(r, A, ent3[j], box, ev + 4);
        f[3 * p] = (float)acc.x;
        f[3 * p + 1] = (float)acc.y;
        f[3 * p + 2] = (float)acc.z;
    }
    if (partials) block_reduce_store<8>(ev, partials + 8 * blockIdx.x);
}

struct DihArgs {
    const int32_t *a, *b, *c, *d, *dtype, *last;
    const float *coeff;
    int dipole_flag;
};

// One dihedral term t of the caller's arrays (dihedral_core on its four
// positions; dipoles and transfer matrices at rows 12 t and 54 t).
template <typename PosT>
__device__ void dihedral_term(const PosT *__restrict__ r, const DihArgs &D, int64_t t, const Box &box, double *ev,
                              PosT *__restrict__ dipoles, PosT *__restrict__ tmat, V3 (&out)[4]) {
    dihedral_core(load3(r, D.a[t]), load3(r, D.b[t]), load3(r, D.c[t]), load3(r, D.d[t]), D.coeff + 30 * t,
                  D.dtype[t], D.last[t], box, D.dipole_flag, ev, dipoles + 12 * t, tmat + 54 * t, out);
}

// One thread per dihedral: its dipoles and transfer matrices alone.
__global__ void __launch_bounds__(kBlock) dihedral_dipole_kernel(const double *__restrict__ r, int64_t m, DihArgs D,
                                                                  Box box, double *__restrict__ dipoles,
                                                                  double *__restrict__ tmat) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride)
        dihedral_dipoles_core(load3(r, D.a[t]), load3(r, D.b[t]), load3(r, D.c[t]), load3(r, D.d[t]), D.dtype[t],
                              D.last[t], box, dipoles + 12 * t, tmat + 54 * t);
}

// One thread per dihedral: its four particle forces into slots (term-major,
// 4 roles x 3 doubles), energies and clamps into per-block partials.
template <typename PosT>
__global__ void __launch_bounds__(kBlock) dihedral_term_kernel(const PosT *__restrict__ r, int64_t m, DihArgs D,
                                                                Box box, double *__restrict__ slots,
                                                                PosT *__restrict__ dipoles, PosT *__restrict__ tmat,
                                                                double *__restrict__ partials) {
    double ev[2] = {0, 0};
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < m; t += stride) {
        V3 out[4];
        dihedral_term(r, D, t, box, ev, dipoles, tmat, out);
        double *sl = slots + 12 * t;
        for (int q = 0; q < 4; ++q) {
            sl[3 * q] = out[q].x;
            sl[3 * q + 1] = out[q].y;
            sl[3 * q + 2] = out[q].z;
        }
    }
    if (partials) block_reduce_store<2>(ev, partials + 2 * blockIdx.x);
}

// Per particle: the slots of its dihedral entries summed in term order (the
// reference's per-particle order), every particle written (zero if none).
template <typename PosT>
__global__ void __launch_bounds__(kBlock) dihedral_slot_gather_kernel(PosT *__restrict__ f, int64_t n,
                                                                       const int *__restrict__ off,
                                                                       const uint32_t *__restrict__ ent,
                                                                       const double *__restrict__ slots) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        V3 acc = v3(0, 0, 0);
        for (int j = off[p]; j < off[p + 1]; ++j) {
            const uint32_t e = ent[j];
            const double *sl = slots + 12 * (int64_t)(e >> 2) + 3 * (e & 3);
            acc = acc + v3(sl[0], sl[1], sl[2]);
        }
        f[3 * p] = (PosT)acc.x;
        f[3 * p + 1] = (PosT)acc.y;
        f[3 * p + 2] = (PosT)acc.z;
    }
}

// force.py:855-877 as a gather over the dihedral lists
template <typename PosT>
__global__ void __launch_bounds__(kBlock) redistribution_kernel(PosT *__restrict__ out, int64_t n,
                                                                 const int *__restrict__ off,
                                                                 const uint32_t *__restrict__ ent,
                                                                 const PosT *__restrict__ fdip,
                                                                 const PosT *__restrict__ tmat,
                                                                 const int32_t *__restrict__ dtype,
                                                                 const int32_t *__restrict__ last) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        double acc[3] = {0, 0, 0};
        for (int j = off[p]; j < off[p + 1]; ++j) {
            const int64_t t = ent[j] >> 2;
            const int role = ent[j] & 3;
            if (dtype[t] != 1) continue;
            const int npairs = last[t] == 1 ? 2 : 1;
            for (int pair = 0; pair < npairs; ++pair) {
                // pair 0 acts on (a,b,c) = roles (0,1,2); pair 1 on (b,c,d) = roles (1,2,3)
                const int x = role - pair;
                if (x < 0 || x > 2) continue;
                const PosT *f0 = fdip + 12 * t + 6 * pair;
                const PosT *f1 = f0 + 3;
                const PosT *M = tmat + 54 * t + 27 * pair + 9 * x;
                double diff[3], sum[3];
                for (int q = 0; q < 3; ++q) {
                    diff[q] = (double)f0[q] - (double)f1[q];
                    sum[q] = (double)f0[q] + (double)f1[q];
                }
                for (int i = 0; i < 3; ++i) {
                    double s = (double)M[3 * i] * diff[0] + (double)M[3 * i + 1] * diff[1] + (double)M[3 * i + 2] * diff[2];
                    if (x > 0) s += 0.5 * sum[i];
                    acc[i] += s;
                }
            }
        }
        out[3 * p] = (PosT)acc[0];
        out[3 * p + 1] = (PosT)acc[1];
        out[3 * p + 2] = (PosT)acc[2];
    }
}

// Workspace of the stateless f2py-style entry points (rebuilt every call, as
// the Fortran kernels recompute everything; kept for the process lifetime).
struct Workspace {
    std::mutex mu;
    Csr csr;
    DevBuf slots;  // per-term force contributions (dihedrals: 4 roles x 3 doubles)
    DevBuf partials, result;
    HostBuf host;
};
static Workspace &workspace() {
    static Workspace *w = new Workspace();  // never freed: outlives the HIP runtime teardown
    return *w;
}

static int finish_reduce(Workspace &w, int nblocks, int nv, double *out, hipStream_t s) {
    HYMD_TRY(w.result.reserve(sizeof(double) * nv));
    HYMD_TRY(w.host.reserve(sizeof(double) * nv));
    HYMD_TRY(reduce_partials((const double *)w.partials.ptr, nblocks, nv, (double *)w.result.ptr, s));
    HYMD_HIP(hipMemcpyAsync(w.host.ptr, w.result.ptr, sizeof(double) * nv, hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    for (int j = 0; j < nv; ++j) out[j] = ((double *)w.host.ptr)[j];
    return HYMD_OK;
}

static Box make_box(const double box[3]) { return make_box3(box); }

}  // namespace hymd

using namespace hymd;

extern "C" int hymd_cbf(void *f, const void *r, int64_t n, const double box[3], const int32_t *a, const int32_t *b,
                        const double *r0, const double *k, int64_t m, int32_t precision, double *energy,
                        double *bond_pr, void *stream) {
    HYMD_REQUIRE(f && box && n >= 0 && m >= 0 && (r || n == 0), "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && r0 && k), "null term array");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    hipStream_t s = (hipStream_t)stream;
    Workspace &w = workspace();
    std::lock_guard<std::mutex> lock(w.mu);
    TermIdx ti{{a, b, nullptr, nullptr}, 2};
    HYMD_TRY(build_csr(w.csr, ti, m, n, s));
    const int nb = grid_for(n, kBlock, 2048);
    const bool red = energy || bond_pr;
    if (red) HYMD_TRY(w.partials.reserve(sizeof(double) * 4 * nb));
    double *part = red ? (double *)w.partials.ptr : nullptr;
    BondArgs B{a, b, r0, k};
    Box bx = make_box(box);
    if (precision == HYMD_F32)
        bond_gather_kernel<float, float><<<nb, kBlock, 0, s>>>((const float *)r, (float *)f, n, w.csr.off(),
                                                                w.csr.ent(), B, bx, part);
    else
        bond_gather_kernel<double, double><<<nb, kBlock, 0, s>>>((const double *)r, (double *)f, n, w.csr.off(),
                                                                  w.csr.ent(), B, bx, part);
    HYMD_HIP(hipGetLastError());
    if (red) {
        double out[4];
        HYMD_TRY(finish_reduce(w, nb, 4, out, s));
        if (energy) *energy = out[0];
        if (bond_pr)
            for (int d = 0; d < 3; ++d) bond_pr[d] = out[1 + d];
    }
    return HYMD_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free bonds: hymd_cbf's kernels with every scratch buffer (the
// per-particle term lists, the energy / virial partials and their sums) in
// one file-scope device array.  Each device holds its own copy of the array
// for the life of the process; its address is looked up once per device and
// kept.  All calls on one device share it, so calls that overlap on a device
// race.
// ---------------------------------------------------------------------------
constexpr int64_t kScratchMaxParticles = HYMD_NOALLOC_MAX_PARTICLES;
constexpr int64_t kScratchMaxBonds = HYMD_NOALLOC_MAX_BONDS;
constexpr int kScratchMaxBlocks = 2048;  // the grid cap of the gather, so the partial rows of any n fit
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kScrCounts = 0;
constexpr size_t kScrOffsets = kScrCounts + align256(sizeof(int) * kScratchMaxParticles);
constexpr size_t kScrEntries = kScrOffsets + align256(sizeof(int) * (kScratchMaxParticles + 1));
constexpr size_t kScrTiles = kScrEntries + align256(sizeof(uint32_t) * 2 * kScratchMaxBonds);
constexpr size_t kScrPartials =
    kScrTiles + align256(sizeof(int) * ((kScratchMaxParticles + 1023) / 1024 + 2));
constexpr size_t kScrResult = kScrPartials + align256(sizeof(double) * 4 * kScratchMaxBlocks);
constexpr size_t kScratchBytes = kScrResult + align256(sizeof(double) * 4);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static int scratch_base(unsigned char **out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned char *ptr[kMaxDevices] = {};
    int dev = -1;
    HYMD_HIP(hipGetDevice(&dev));
    HYMD_REQUIRE(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        HYMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        HYMD_REQUIRE(p != nullptr, "no device address for the bond scratch array");
        ptr[dev] = (unsigned char *)p;
    }
    *out = ptr[dev];
    return HYMD_OK;
}

extern "C" int hymd_cbf_noalloc(void *f, const void *r, int64_t n, const double box[3], const int32_t *a,
                                const int32_t *b, const double *r0, const double *k, int64_t m, int32_t precision,
                                double *energy, double *bond_pr, void *stream) {
    HYMD_REQUIRE(f && box && n >= 0 && m >= 0 && (r || n == 0), "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && r0 && k), "null term array");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    HYMD_REQUIRE(n <= kScratchMaxParticles, "%lld particles exceed the allocation-free limit of %lld",
                 (long long)n, (long long)kScratchMaxParticles);
    HYMD_REQUIRE(m <= kScratchMaxBonds, "%lld bonds exceed the allocation-free limit of %lld", (long long)m,
                 (long long)kScratchMaxBonds);
    hipStream_t s = (hipStream_t)stream;
    unsigned char *base = nullptr;
    HYMD_TRY(scratch_base(&base));
    HYMD_REQUIRE(scan_tiles(n) + 2 <= (int64_t)((kScrPartials - kScrTiles) / sizeof(int)), "scan tile scratch");
    int *counts = (int *)(base + kScrCounts), *off = (int *)(base + kScrOffsets), *tiles = (int *)(base + kScrTiles);
    uint32_t *entries = (uint32_t *)(base + kScrEntries);
    double *partials = (double *)(base + kScrPartials), *result = (double *)(base + kScrResult);
    TermIdx ti{{a, b, nullptr, nullptr}, 2};
    HYMD_TRY(build_csr_into(counts, off, entries, tiles, ti, m, n, s));
    const int nb = grid_for(n, kBlock, kScratchMaxBlocks);
    const bool red = energy || bond_pr;
    double *part = red ? partials : nullptr;
    BondArgs B{a, b, r0, k};
    Box bx = make_box(box);
    if (precision == HYMD_F32)
        bond_gather_kernel<float, float><<<nb, kBlock, 0, s>>>((const float *)r, (float *)f, n, off, entries, B, bx,
                                                                part);
    else
        bond_gather_kernel<double, double><<<nb, kBlock, 0, s>>>((const double *)r, (double *)f, n, off, entries, B,
                                                                  bx, part);
    HYMD_HIP(hipGetLastError());
    if (red) {
        double out[4];
        HYMD_TRY(reduce_partials(partials, nb, 4, result, s));
        HYMD_HIP(hipMemcpyAsync(out, result, sizeof(out), hipMemcpyDeviceToHost, s));
        HYMD_HIP(hipStreamSynchronize(s));
        if (energy) *energy = out[0];
        if (bond_pr)
            for (int d = 0; d < 3; ++d) bond_pr[d] = out[1 + d];
    }
    return HYMD_OK;
}

extern "C" int hymd_caf(void *f, const void *r, int64_t n, const double box[3], const int32_t *a, const int32_t *b,
                        const int32_t *c, const double *t0, const double *k, int64_t m, int32_t precision,
                        double *energy, double *angle_pr, void *stream) {
    HYMD_REQUIRE(f && box && n >= 0 && m >= 0 && (r || n == 0), "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && c && t0 && k), "null term array");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    hipStream_t s = (hipStream_t)stream;
    Workspace &w = workspace();
    std::lock_guard<std::mutex> lock(w.mu);
    TermIdx ti{{a, b, c, nullptr}, 3};
    HYMD_TRY(build_csr(w.csr, ti, m, n, s));
    const int nb = grid_for(n, kBlock, 2048);
    const bool red = energy || angle_pr;
    if (red) HYMD_TRY(w.partials.reserve(sizeof(double) * 4 * nb));
    double *part = red ? (double *)w.partials.ptr : nullptr;
    AngleArgs A{a, b, c, t0, k};
    Box bx = make_box(box);
    if (precision == HYMD_F32)
        angle_gather_kernel<float, float><<<nb, kBlock, 0, s>>>((const float *)r, (float *)f, n, w.csr.off(),
                                                                 w.csr.ent(), A, bx, part);
    else
        angle_gather_kernel<double, double><<<nb, kBlock, 0, s>>>((const double *)r, (double *)f, n, w.csr.off(),
                                                                   w.csr.ent(), A, bx, part);
    HYMD_HIP(hipGetLastError());
    if (red) {
        double out[4];
        HYMD_TRY(finish_reduce(w, nb, 4, out, s));
        if (energy) *energy = out[0];
        if (angle_pr)
            for (int d = 0; d < 3; ++d) angle_pr[d] = out[1 + d];
    }
    return HYMD_OK;
}

extern "C" int hymd_cdf(void *f, const void *r, void *dipoles, void *transfer_matrix, int64_t n, const double box[3],
                        const int32_t *a, const int32_t *b, const int32_t *c, const int32_t *d, const float *coeff,
                        const int32_t *dtype, const int32_t *bb_index, int64_t m, int32_t dipole_flag,
                        int32_t precision, double *energy, int64_t *n_clamped, void *stream) {
    HYMD_REQUIRE(f && box && n >= 0 && m >= 0 && (r || n == 0), "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && c && d && coeff && dtype && bb_index && dipoles && transfer_matrix),
                 "null term array");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    hipStream_t s = (hipStream_t)stream;
    Workspace &w = workspace();
    std::lock_guard<std::mutex> lock(w.mu);
    TermIdx ti{{a, b, c, d}, 4};
    HYMD_TRY(build_csr(w.csr, ti, m, n, s));
    HYMD_TRY(w.slots.reserve(sizeof(double) * 12 * (size_t)std::max<int64_t>(m, 1)));
    const int nb = grid_for(m, kBlock, 2048);
    const bool red = energy || n_clamped;
    if (red) HYMD_TRY(w.partials.reserve(sizeof(double) * 2 * nb));
    double *part = red ? (double *)w.partials.ptr : nullptr;
    DihArgs D{a, b, c, d, dtype, bb_index, coeff, dipole_flag};
    Box bx = make_box(box);
    double *slots = (double *)w.slots.ptr;
    if (precision == HYMD_F32) {
        if (m > 0)
            dihedral_term_kernel<float><<<nb, kBlock, 0, s>>>((const float *)r, m, D, bx, slots, (float *)dipoles,
                                                              (float *)transfer_matrix, part);
        if (n > 0)
            dihedral_slot_gather_kernel<float><<<grid_for(n), kBlock, 0, s>>>((float *)f, n, w.csr.off(),
                                                                            w.csr.ent(), slots);
    } else {
        if (m > 0)
            dihedral_term_kernel<double><<<nb, kBlock, 0, s>>>((const double *)r, m, D, bx, slots, (double *)dipoles,
                                                               (double *)transfer_matrix, part);
        if (n > 0)
            dihedral_slot_gather_kernel<double><<<grid_for(n), kBlock, 0, s>>>((double *)f, n, w.csr.off(),
                                                                             w.csr.ent(), slots);
    }
    HYMD_HIP(hipGetLastError());
    if (red) {
        double out[2] = {0.0, 0.0};
        if (m > 0) HYMD_TRY(finish_reduce(w, nb, 2, out, s));
        if (energy) *energy = out[0];
        if (n_clamped) *n_clamped = (int64_t)out[1];
    }
    return HYMD_OK;
}

extern "C" int hymd_dipole_redistribution(void *f_on_bead, int64_t n, const void *f_dipoles,
                                          const void *transfer_matrix, const int32_t *a, const int32_t *b,
                                          const int32_t *c, const int32_t *d, const int32_t *dtype,
                                          const int32_t *last, int64_t m, int32_t precision, void *stream) {
    HYMD_REQUIRE(f_on_bead && n >= 0 && m >= 0, "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && c && d && dtype && last && f_dipoles && transfer_matrix), "null term array");
    HYMD_REQUIRE(precision == HYMD_F32 || precision == HYMD_F64, "precision must be HYMD_F32 or HYMD_F64");
    hipStream_t s = (hipStream_t)stream;
    Workspace &w = workspace();
    std::lock_guard<std::mutex> lock(w.mu);
    TermIdx ti{{a, b, c, d}, 4};
    HYMD_TRY(build_csr(w.csr, ti, m, n, s));
    if (precision == HYMD_F32)
        redistribution_kernel<float><<<grid_for(n), kBlock, 0, s>>>((float *)f_on_bead, n, w.csr.off(), w.csr.ent(),
                                                                    (const float *)f_dipoles,
                                                                    (const float *)transfer_matrix, dtype, last);
    else
        redistribution_kernel<double><<<grid_for(n), kBlock, 0, s>>>((double *)f_on_bead, n, w.csr.off(),
                                                                     w.csr.ent(), (const double *)f_dipoles,
                                                                     (const double *)transfer_matrix, dtype, last);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// ---------------------------------------------------------------------------
// persistent topology
// ---------------------------------------------------------------------------
struct hymd_topology {
    int64_t n = 0, m2 = 0, m3 = 0;
    Csr c2, c3;
    BondArgs B{};
    AngleArgs A{};
    DevBuf partials, result;
    HostBuf host;
    int nblocks = 1;
};

extern "C" int hymd_topology_create(hymd_topology **out, int64_t n, int64_t m2, const int32_t *a2, const int32_t *b2,
                                    const double *r0, const double *k2, int64_t m3, const int32_t *a3,
                                    const int32_t *b3, const int32_t *c3, const double *t0, const double *k3,
                                    void *stream) {
    HYMD_REQUIRE(out && n >= 0 && m2 >= 0 && m3 >= 0, "bad argument");
    HYMD_REQUIRE(m2 == 0 || (a2 && b2 && r0 && k2), "null bond array");
    HYMD_REQUIRE(m3 == 0 || (a3 && b3 && c3 && t0 && k3), "null angle array");
    hipStream_t s = (hipStream_t)stream;
    hymd_topology *t = new hymd_topology();
    t->n = n;
    t->m2 = m2;
    t->m3 = m3;
    t->B = BondArgs{a2, b2, r0, k2};
    t->A = AngleArgs{a3, b3, c3, t0, k3};
    int st = HYMD_OK;
    if (m2 > 0) {
        TermIdx ti{{a2, b2, nullptr, nullptr}, 2};
        st = build_csr(t->c2, ti, m2, n, s);
    }
    if (st == HYMD_OK && m3 > 0) {
        TermIdx ti{{a3, b3, c3, nullptr}, 3};
        st = build_csr(t->c3, ti, m3, n, s);
    }
    t->nblocks = grid_for(n, kBlock, 2048);
    if (st == HYMD_OK) st = t->partials.reserve(sizeof(double) * 8 * t->nblocks);
    if (st == HYMD_OK) st = t->result.reserve(sizeof(double) * 8);
    if (st == HYMD_OK) st = t->host.reserve(sizeof(double) * 8);
    if (st != HYMD_OK) {
        delete t;
        return st;
    }
    *out = t;
    return HYMD_OK;
}

namespace hymd {
TopoTerms topology_terms(const hymd_topology *t) {
    TopoTerms v;
    v.n = t->n;
    v.m2 = t->m2;
    v.m3 = t->m3;
    v.a2 = t->B.a;
    v.b2 = t->B.b;
    v.r0 = t->B.r0;
    v.k2 = t->B.k;
    v.a3 = t->A.a;
    v.b3 = t->A.b;
    v.c3 = t->A.c;
    v.t0 = t->A.t0;
    v.k3 = t->A.k;
    return v;
}
}  // namespace hymd

extern "C" int hymd_topology_destroy(hymd_topology *top) {
    delete top;
    return HYMD_OK;
}

extern "C" int hymd_topology_forces(hymd_topology *top, const double *r, float *f, const double box[3],
                                    double *energies, double *virials, void *stream) {
    HYMD_REQUIRE(top && f && box && (r || top->n == 0), "bad argument");
    hipStream_t s = (hipStream_t)stream;
    const bool red = energies || virials;
    double *part = red ? (double *)top->partials.ptr : nullptr;
    if (top->n > 0)
        topo_gather_kernel<<<top->nblocks, kBlock, 0, s>>>(
            r, f, top->n, top->m2 ? top->c2.off() : nullptr, top->m2 ? top->c2.ent() : nullptr, top->B,
            top->m3 ? top->c3.off() : nullptr, top->m3 ? top->c3.ent() : nullptr, top->A, make_box(box), part);
    HYMD_HIP(hipGetLastError());
    if (red) {
        HYMD_TRY(reduce_partials(part, top->nblocks, 8, (double *)top->result.ptr, s));
        HYMD_HIP(hipMemcpyAsync(top->host.ptr, top->result.ptr, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
        HYMD_HIP(hipStreamSynchronize(s));
        const double *h = (const double *)top->host.ptr;
        if (energies) {
            energies[0] = h[0];
            energies[1] = h[4];
        }
        if (virials)
            for (int d = 0; d < 3; ++d) {
                virials[d] = h[1 + d];
                virials[3 + d] = h[5 + d];
            }
    }
    return HYMD_OK;
}

// ---------------------------------------------------------------------------
// Persistent dihedral term list: per-particle lists built once, term arrays
// referenced (not copied), FP64 forces without per-call synchronisation.
// ---------------------------------------------------------------------------
struct hymd_dihedrals {
    int64_t n = 0, m = 0;
    Csr csr;
    DihArgs D{};
    DevBuf slots, partials, result;
    HostBuf host;
    int nblocks = 1;
};

namespace hymd {
DihTerms dihedral_terms(const hymd_dihedrals *h) {
    DihTerms v;
    v.m = h->m;
    v.a = h->D.a;
    v.b = h->D.b;
    v.c = h->D.c;
    v.d = h->D.d;
    v.dtype = h->D.dtype;
    v.last = h->D.last;
    v.coeff = h->D.coeff;
    return v;
}
}  // namespace hymd

extern "C" int hymd_dihedrals_create(hymd_dihedrals **out, int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                                     const int32_t *c, const int32_t *d, const float *coeff, const int32_t *dtype,
                                     const int32_t *last, void *stream) {
    HYMD_REQUIRE(out && n >= 0 && m >= 0, "bad argument");
    HYMD_REQUIRE(m == 0 || (a && b && c && d && coeff && dtype && last), "null term array");
    hipStream_t s = (hipStream_t)stream;
    std::unique_ptr<hymd_dihedrals> h(new hymd_dihedrals());
    h->n = n;
    h->m = m;
    h->D = DihArgs{a, b, c, d, dtype, last, coeff, 0};
    TermIdx ti{{a, b, c, d}, 4};
    HYMD_TRY(build_csr(h->csr, ti, m, n, s));  // validates the indices (synchronises once)
    h->nblocks = grid_for(m, kBlock, 2048);
    HYMD_TRY(h->slots.reserve(sizeof(double) * 12 * (size_t)std::max<int64_t>(m, 1)));
    HYMD_TRY(h->partials.reserve(sizeof(double) * 2 * h->nblocks));
    HYMD_TRY(h->result.reserve(sizeof(double) * 2));
    HYMD_TRY(h->host.reserve(sizeof(double) * 2));
    *out = h.release();
    return HYMD_OK;
}

extern "C" int hymd_dihedrals_destroy(hymd_dihedrals *h) {
    delete h;
    return HYMD_OK;
}

extern "C" int hymd_dihedrals_forces(hymd_dihedrals *h, const double *x, double *f, double *dipoles, double *tmat,
                                     const double box[3], int32_t dipole_flag, double *energy, void *stream) {
    HYMD_REQUIRE(h && box && f && (h->n == 0 || x) && (h->m == 0 || (dipoles && tmat)), "bad argument");
    hipStream_t s = (hipStream_t)stream;
    DihArgs D = h->D;
    D.dipole_flag = dipole_flag;
    Box bx = make_box(box);
    double *part = energy ? (double *)h->partials.ptr : nullptr;
    if (h->m > 0)
        dihedral_term_kernel<double><<<h->nblocks, kBlock, 0, s>>>(x, h->m, D, bx, (double *)h->slots.ptr, dipoles,
                                                                   tmat, part);
    if (h->n > 0)
        dihedral_slot_gather_kernel<double><<<grid_for(h->n), kBlock, 0, s>>>(f, h->n, h->csr.off(), h->csr.ent(),
                                                                            (const double *)h->slots.ptr);
    HYMD_HIP(hipGetLastError());
    if (energy) {
        *energy = 0.0;
        if (h->m > 0) {
            HYMD_TRY(reduce_partials((const double *)h->partials.ptr, h->nblocks, 2, (double *)h->result.ptr, s));
            HYMD_HIP(hipMemcpyAsync(h->host.ptr, h->result.ptr, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
            HYMD_HIP(hipStreamSynchronize(s));
            *energy = ((const double *)h->host.ptr)[0];
        }
    }
    return HYMD_OK;
}

extern "C" int hymd_dihedrals_dipoles(hymd_dihedrals *h, const double *x, double *dipoles, double *tmat,
                                      const double box[3], void *stream) {
    HYMD_REQUIRE(h && box && (h->m == 0 || (x && dipoles && tmat)), "bad argument");
    if (h->m > 0)
        dihedral_dipole_kernel<<<h->nblocks, kBlock, 0, (hipStream_t)stream>>>(x, h->m, h->D, make_box(box), dipoles,
                                                                               tmat);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

extern "C" int hymd_dihedrals_redistribute(hymd_dihedrals *h, const double *f_dipoles, const double *tmat, double *f,
                                           void *stream) {
    HYMD_REQUIRE(h && f && (h->m == 0 || (f_dipoles && tmat)), "bad argument");
    hipStream_t s = (hipStream_t)stream;
    if (h->n > 0)
        redistribution_kernel<double><<<grid_for(h->n), kBlock, 0, s>>>(f, h->n, h->csr.off(), h->csr.ent(), f_dipoles,
                                                                         tmat, h->D.dtype, h->D.last);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}
