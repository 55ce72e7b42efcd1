// prepare_bonds (force.py:573-728 over prepare_bonds_old :325-505) on the
// device: the bonded term arrays from the per-particle bond table.
//
// One thread per molecule (molecules are independent; their terms are
// concatenated in molecule order), two passes over the same traversal:
//   graph  build the molecule's graph in networkx.Graph insertion order
//          (nodes in order of first appearance, neighbours in order of first
//          edge insertion) in per-particle scratch
//   count  walk it, count bonds / angles / dihedrals
//   scan   exclusive scans of the three counts -> each molecule's offsets
//   fill   walk it again, write the terms at the offsets
// The walk (per source node s, in node order):
//   bonds      level 1 of the breadth-first search from s (neighbour order),
//              j with a larger global index than s, every matching type
//   angles     level 2 (for each level-1 node v in order, its unseen
//              neighbours w, reached through v): (s, v, w), gid(w) > gid(s)
//   dihedrals  simple paths s-a-b-c nested in neighbour order (find_all_paths,
//              force.py:266-276), every type whose names match forward; the
//              reference skips a term whose reversed copy with the same type
//              object is already stored: the reversed path c-b-a-s was walked
//              before iff c precedes s in node order, and it stored every type
//              it matches forward, in type order; the reference's list
//              comparison meets the first of those, and raises ValueError when
//              that one is another type (numpy array truth value)
// Per molecule the last stored dihedral of type 1 gets last = 1.
#include <vector>

#include "common.h"

namespace hymd {

constexpr int kTopoMaxTypes = 256;  // bond / angle / dihedral type entries each

enum TopoError : unsigned { kTopoOk = 0, kTopoKeyError = 1, kTopoAmbiguous = 2, kTopoDegree = 3 };

struct TopoTables {
    int n2, n3, n4;
    const int *names2;      // [n2][2]
    const int *names3;      // [n3][3]
    const int *names4;      // [n4][4]
    const int *kind4;       // [n4] dih_type
    const double *par2;     // [n2][2] r0, k
    const double *par3;     // [n3][2] theta0 (rad), k
    const double *coef4;    // [n4][30]
};

struct TopoGraph {
    int64_t n_mol;
    const int64_t *mol_start;  // [n_mol + 1]
    const int64_t *gid;        // [n] global index (node key)
    const int64_t *row;        // [n] local index written out, or NULL (position)
    const int *name;           // [n]
    const int64_t *bonds;      // [n][max_bonds] global indices, -1 = none
    int max_bonds;
    int cap;                   // adjacency capacity per node
    // scratch, indexed from the molecule's first row
    int *slot_of;              // member -> node slot
    int *member_of;            // node slot -> member
    int *deg;                  // [slot]
    int *adj;                  // [slot][cap] neighbour slots
    int *mark;                 // [slot] BFS stamps
    unsigned long long *err;   // min over (molecule << 3 | kind)
};

__device__ __forceinline__ void topo_fail(const TopoGraph &g, int64_t mol, unsigned kind) {
    atomicMin(g.err, ((unsigned long long)mol << 3) | kind);
}

// member position of global index `key` in the molecule's rows [s0, s0 + m), -1 if none
__device__ int find_member(const int64_t *__restrict__ gid, int64_t s0, int m, int64_t key, bool consecutive) {
    if (consecutive) {
        const int64_t d = key - gid[s0];
        return (d >= 0 && d < m) ? (int)d : -1;
    }
    for (int k = 0; k < m; ++k)
        if (gid[s0 + k] == key) return k;
    return -1;
}

__global__ void __launch_bounds__(kBlock) topo_graph_kernel(TopoGraph g) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t mol = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; mol < g.n_mol; mol += stride) {
        const int64_t s0 = g.mol_start[mol];
        const int m = (int)(g.mol_start[mol + 1] - s0);
        bool consecutive = true;
        for (int k = 0; k < m; ++k) {
            g.slot_of[s0 + k] = -1;
            consecutive = consecutive && g.gid[s0 + k] == g.gid[s0] + k;
        }
        int nodes = 0;
        bool ok = true;
        for (int k = 0; k < m && ok; ++k) {
            if (g.slot_of[s0 + k] < 0) {  // add_node
                g.slot_of[s0 + k] = nodes;
                g.member_of[s0 + nodes] = k;
                g.deg[s0 + nodes] = 0;
                ++nodes;
            }
            const int u = g.slot_of[s0 + k];
            for (int c = 0; c < g.max_bonds; ++c) {
                const int64_t key = g.bonds[(s0 + k) * g.max_bonds + c];
                if (key == -1) continue;
                const int p = find_member(g.gid, s0, m, key, consecutive);
                if (p < 0) {  // the partner has no node attributes in the reference
                    topo_fail(g, mol, kTopoKeyError);
                    ok = false;
                    break;
                }
                if (g.slot_of[s0 + p] < 0) {  // add_edge adds the partner node
                    g.slot_of[s0 + p] = nodes;
                    g.member_of[s0 + nodes] = p;
                    g.deg[s0 + nodes] = 0;
                    ++nodes;
                }
                const int v = g.slot_of[s0 + p];
                if (u == v) continue;  // a self-loop: no traversal of the reference can use it
                int *au = g.adj + (s0 + u) * g.cap;
                const int du = g.deg[s0 + u];
                bool seen = false;
                for (int q = 0; q < du; ++q) seen = seen || au[q] == v;
                if (seen) continue;
                const int dv = g.deg[s0 + v];
                if (du >= g.cap || dv >= g.cap) {
                    topo_fail(g, mol, kTopoDegree);
                    ok = false;
                    break;
                }
                au[du] = v;
                g.adj[(s0 + v) * g.cap + dv] = u;
                g.deg[s0 + u] = du + 1;
                g.deg[s0 + v] = dv + 1;
            }
        }
    }
}

struct TopoOut {
    const int *off2, *off3, *off4;  // exclusive offsets per molecule
    int *cnt2, *cnt3, *cnt4;        // counts per molecule (count pass)
    int *a2, *b2, *t2;
    double *eq2, *k2;
    int *a3, *b3, *c3, *t3;
    double *eq3, *k3;
    int *a4, *b4, *c4, *d4, *t4, *kind4, *last4;
    double *coef4;                  // [D][30]
};

template <bool FILL>
__global__ void __launch_bounds__(kBlock) topo_walk_kernel(TopoGraph g, TopoTables t, TopoOut o) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t mol = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; mol < g.n_mol; mol += stride) {
        const int64_t s0 = g.mol_start[mol];
        const int m = (int)(g.mol_start[mol + 1] - s0);
        if (*g.err != ~0ull && (int64_t)(*g.err >> 3) <= mol) continue;  // this or an earlier molecule failed
        const int *adj = g.adj + s0 * g.cap;
        const int *deg = g.deg + s0;
        int *mark = g.mark + s0;
        auto member = [&](int slot) { return g.member_of[s0 + slot]; };
        auto key = [&](int slot) { return g.gid[s0 + member(slot)]; };
        auto nm = [&](int slot) { return g.name[s0 + member(slot)]; };
        auto out_row = [&](int slot) {
            const int64_t r = s0 + member(slot);
            return (int)(g.row ? g.row[r] : r);
        };
        for (int s = 0; s < m; ++s) mark[s] = 0;
        int n2 = 0, n3 = 0, n4 = 0, bb = -1;
        const int64_t o2 = FILL ? o.off2[mol] : 0, o3 = FILL ? o.off3[mol] : 0, o4 = FILL ? o.off4[mol] : 0;
        bool ok = true;
        for (int s = 0; s < m && ok; ++s) {
            const int stamp = s + 1;
            const int64_t ks = key(s);
            const int ns = nm(s);
            mark[s] = stamp;
            // level 1: bonds
            for (int q = 0; q < deg[s]; ++q) {
                const int j = adj[s * g.cap + q];
                mark[j] = stamp;
                if (key(j) <= ks) continue;
                const int nj = nm(j);
                for (int e = 0; e < t.n2; ++e) {
                    const int e0 = t.names2[2 * e], e1 = t.names2[2 * e + 1];
                    if (!((ns == e0 && nj == e1) || (ns == e1 && nj == e0))) continue;
                    if (FILL) {
                        const int64_t w = o2 + n2;
                        o.a2[w] = out_row(s);
                        o.b2[w] = out_row(j);
                        o.eq2[w] = t.par2[2 * e];
                        o.k2[w] = t.par2[2 * e + 1];
                        if (o.t2) o.t2[w] = e;
                    }
                    ++n2;
                }
            }
            // level 2: angles (s, v, w)
            for (int q = 0; q < deg[s]; ++q) {
                const int v = adj[s * g.cap + q];
                const int nv = nm(v);
                for (int r = 0; r < deg[v]; ++r) {
                    const int w = adj[v * g.cap + r];
                    if (mark[w] == stamp) continue;
                    mark[w] = stamp;
                    if (key(w) <= ks) continue;
                    const int nw = nm(w);
                    for (int e = 0; e < t.n3; ++e) {
                        const int e0 = t.names3[3 * e], e1 = t.names3[3 * e + 1], e2 = t.names3[3 * e + 2];
                        if (nv != e1 || !((ns == e0 && nw == e2) || (ns == e2 && nw == e0))) continue;
                        if (FILL) {
                            const int64_t w3 = o3 + n3;
                            o.a3[w3] = out_row(s);
                            o.b3[w3] = out_row(v);
                            o.c3[w3] = out_row(w);
                            o.eq3[w3] = t.par3[2 * e];
                            o.k3[w3] = t.par3[2 * e + 1];
                            if (o.t3) o.t3[w3] = e;
                        }
                        ++n3;
                    }
                }
            }
            // dihedrals: simple paths s-a-b-c
            if (t.n4 == 0) continue;
            for (int qa = 0; qa < deg[s] && ok; ++qa) {
                const int a = adj[s * g.cap + qa];
                const int na = nm(a);
                for (int qb = 0; qb < deg[a] && ok; ++qb) {
                    const int b = adj[a * g.cap + qb];
                    if (b == s) continue;
                    const int nb = nm(b);
                    for (int qc = 0; qc < deg[b] && ok; ++qc) {
                        const int c = adj[b * g.cap + qc];
                        if (c == a || c == s) continue;
                        const int nc = nm(c);
                        for (int e = 0; e < t.n4; ++e) {
                            const int *en = t.names4 + 4 * e;
                            if (ns != en[0] || na != en[1] || nb != en[2] || nc != en[3]) continue;
                            if (c < s) {  // the reversed path stored its types already
                                int first = -1;
                                for (int f = 0; f < t.n4 && first < 0; ++f) {
                                    const int *fn = t.names4 + 4 * f;
                                    if (nc == fn[0] && nb == fn[1] && na == fn[2] && ns == fn[3]) first = f;
                                }
                                if (first == e) continue;
                                if (first >= 0) {
                                    topo_fail(g, mol, kTopoAmbiguous);
                                    ok = false;
                                    break;
                                }
                            }
                            if (FILL) {
                                const int64_t w4 = o4 + n4;
                                o.a4[w4] = out_row(s);
                                o.b4[w4] = out_row(a);
                                o.c4[w4] = out_row(b);
                                o.d4[w4] = out_row(c);
                                o.kind4[w4] = t.kind4[e];
                                o.last4[w4] = 0;
                                if (o.t4) o.t4[w4] = e;
                                for (int q = 0; q < 30; ++q) o.coef4[w4 * 30 + q] = t.coef4[30 * e + q];
                            }
                            if (t.kind4[e] == 1) bb = n4;
                            ++n4;
                        }
                    }
                }
            }
        }
        if (!FILL) {
            o.cnt2[mol] = ok ? n2 : 0;
            o.cnt3[mol] = ok ? n3 : 0;
            o.cnt4[mol] = ok ? n4 : 0;
        } else if (bb >= 0) {
            o.last4[o4 + bb] = 1;
        }
    }
}

__global__ void topo_err_init_kernel(unsigned long long *err) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *err = ~0ull;
}

}  // namespace hymd

using namespace hymd;

struct hymd_bond_builder {
    hipStream_t stream = nullptr;
    int64_t n = 0, n_mol = 0;
    TopoGraph g{};
    TopoTables t{};
    DevBuf scratch, tables, counts, offsets, tiles, errbuf;
    int64_t total[3] = {0, 0, 0};
};

extern "C" int hymd_bond_builder_destroy(hymd_bond_builder *b) {
    if (!b) return HYMD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
    return HYMD_OK;
}

extern "C" int hymd_bond_builder_create(hymd_bond_builder **out, int64_t n, int64_t n_mol, const int64_t *mol_start,
                                        const int64_t *index, const int64_t *row, const int32_t *name,
                                        const int64_t *bonds, int32_t max_bonds, const int32_t *bond_names,
                                        const double *bond_params, int32_t n_bond_types, const int32_t *angle_names,
                                        const double *angle_params, int32_t n_angle_types,
                                        const int32_t *dihedral_names, const int32_t *dihedral_kind,
                                        const double *dihedral_coeffs, int32_t n_dihedral_types, int64_t counts[3],
                                        void *stream) {
    HYMD_REQUIRE(out && counts, "NULL output");
    *out = nullptr;
    HYMD_REQUIRE(n >= 0 && n_mol >= 0 && max_bonds >= 0, "negative size");
    HYMD_REQUIRE(n_mol == 0 || (mol_start && index && name && (bonds || max_bonds == 0)), "NULL input array");
    HYMD_REQUIRE(n_bond_types >= 0 && n_bond_types <= kTopoMaxTypes && n_angle_types >= 0 &&
                     n_angle_types <= kTopoMaxTypes && n_dihedral_types >= 0 && n_dihedral_types <= kTopoMaxTypes,
                 "at most %d types of each kind", kTopoMaxTypes);
    HYMD_REQUIRE((n_bond_types == 0 || (bond_names && bond_params)) &&
                     (n_angle_types == 0 || (angle_names && angle_params)) &&
                     (n_dihedral_types == 0 || (dihedral_names && dihedral_kind && dihedral_coeffs)),
                 "NULL type table");
    auto *b = new hymd_bond_builder();
    b->stream = (hipStream_t)stream;
    b->n = n;
    b->n_mol = n_mol;
    const int cap = max_bonds > 0 ? 2 * max_bonds : 1;
    // scratch: slot_of, member_of, deg, mark [n] + adj [n * cap] ints
    const size_t nints = (size_t)n * (4 + cap) + 1;
    int st = b->scratch.reserve(sizeof(int) * nints);
    // tables: names (2 + 3 + 4 + 1 ints per entry), params (2 + 2 + 30 doubles)
    const size_t tab_ints = (size_t)2 * n_bond_types + 3 * n_angle_types + 5 * n_dihedral_types;
    const size_t tab_dbl = (size_t)2 * n_bond_types + 2 * n_angle_types + 30 * n_dihedral_types;
    if (st == HYMD_OK) st = b->tables.reserve(sizeof(double) * (tab_dbl + 1) + sizeof(int) * (tab_ints + 2));
    if (st == HYMD_OK) st = b->counts.reserve(sizeof(int) * 3 * (size_t)(n_mol + 1));
    if (st == HYMD_OK) st = b->offsets.reserve(sizeof(int) * 3 * (size_t)(n_mol + 2));
    if (st == HYMD_OK) st = b->tiles.reserve(sizeof(int) * (size_t)(scan_tiles(n_mol) + 1));
    if (st == HYMD_OK) st = b->errbuf.reserve(sizeof(unsigned long long));
    if (st != HYMD_OK) {
        delete b;
        return st;
    }
    hipStream_t s = b->stream;
    // host tables -> device (one staging copy; create synchronises anyway)
    std::vector<double> hd(tab_dbl + 1, 0.0);
    std::vector<int> hi(tab_ints + 2, 0);
    size_t pd = 0, pi = 0;
    for (int e = 0; e < 2 * n_bond_types; ++e) hd[pd++] = bond_params[e];
    for (int e = 0; e < 2 * n_angle_types; ++e) hd[pd++] = angle_params[e];
    for (int e = 0; e < 30 * n_dihedral_types; ++e) hd[pd++] = dihedral_coeffs[e];
    for (int e = 0; e < 2 * n_bond_types; ++e) hi[pi++] = bond_names[e];
    for (int e = 0; e < 3 * n_angle_types; ++e) hi[pi++] = angle_names[e];
    for (int e = 0; e < 4 * n_dihedral_types; ++e) hi[pi++] = dihedral_names[e];
    for (int e = 0; e < n_dihedral_types; ++e) hi[pi++] = dihedral_kind[e];
    double *td = (double *)b->tables.ptr;
    int *ti = (int *)(td + tab_dbl + 1);
    HYMD_HIP(hipMemcpyAsync(td, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(ti, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice, s));
    TopoTables &t = b->t;
    t.n2 = n_bond_types;
    t.n3 = n_angle_types;
    t.n4 = n_dihedral_types;
    t.par2 = td;
    t.par3 = td + 2 * n_bond_types;
    t.coef4 = td + 2 * n_bond_types + 2 * n_angle_types;
    t.names2 = ti;
    t.names3 = ti + 2 * n_bond_types;
    t.names4 = ti + 2 * n_bond_types + 3 * n_angle_types;
    t.kind4 = ti + 2 * n_bond_types + 3 * n_angle_types + 4 * n_dihedral_types;

    TopoGraph &g = b->g;
    g.n_mol = n_mol;
    g.mol_start = mol_start;
    g.gid = index;
    g.row = row;
    g.name = name;
    g.bonds = bonds;
    g.max_bonds = max_bonds;
    g.cap = cap;
    int *sc = (int *)b->scratch.ptr;
    g.slot_of = sc;
    g.member_of = sc + n;
    g.deg = sc + 2 * n;
    g.mark = sc + 3 * n;
    g.adj = sc + 4 * n;
    g.err = (unsigned long long *)b->errbuf.ptr;

    int *cnt = (int *)b->counts.ptr;
    TopoOut o{};
    o.cnt2 = cnt;
    o.cnt3 = cnt + (n_mol + 1);
    o.cnt4 = cnt + 2 * (n_mol + 1);
    topo_err_init_kernel<<<1, 64, 0, s>>>(g.err);
    HYMD_HIP(hipGetLastError());
    if (n_mol > 0) {
        const int nb = grid_for(n_mol);
        topo_graph_kernel<<<nb, kBlock, 0, s>>>(g);
        HYMD_HIP(hipGetLastError());
        topo_walk_kernel<false><<<nb, kBlock, 0, s>>>(g, t, o);
        HYMD_HIP(hipGetLastError());
        int *off = (int *)b->offsets.ptr;
        for (int k = 0; k < 3; ++k)
            HYMD_TRY(exclusive_scan(cnt + k * (n_mol + 1), off + k * (n_mol + 2), n_mol, (int *)b->tiles.ptr,
                                    nullptr, false, s));
    }
    unsigned long long err = ~0ull;
    HYMD_HIP(hipMemcpyAsync(&err, g.err, sizeof(err), hipMemcpyDeviceToHost, s));
    int tot[3] = {0, 0, 0};
    if (n_mol > 0)
        for (int k = 0; k < 3; ++k)
            HYMD_HIP(hipMemcpyAsync(&tot[k], (int *)b->offsets.ptr + k * (n_mol + 2) + n_mol, sizeof(int),
                                    hipMemcpyDeviceToHost, s));
    HYMD_HIP(hipStreamSynchronize(s));
    if (err != ~0ull) {
        const long long mol = (long long)(err >> 3);
        const unsigned kind = (unsigned)(err & 7);
        delete b;
        if (kind == kTopoKeyError) set_error("KeyError: molecule %lld has a bond partner outside the molecule", mol);
        else if (kind == kTopoAmbiguous)
            set_error("ValueError: molecule %lld: a dihedral's reversed copy is stored with another type "
                      "(the reference's list comparison of their coefficient arrays is ambiguous)", mol);
        else set_error("molecule %lld: a particle has more than %d distinct bond partners", mol, cap);
        return HYMD_ERR_INVALID;
    }
    for (int k = 0; k < 3; ++k) b->total[k] = counts[k] = tot[k];
    *out = b;
    return HYMD_OK;
}

extern "C" int hymd_bond_builder_fill(hymd_bond_builder *b, int32_t *a2, int32_t *b2, double *eq2, double *k2,
                                      int32_t *a3, int32_t *b3, int32_t *c3, double *eq3, double *k3, int32_t *a4,
                                      int32_t *b4, int32_t *c4, int32_t *d4, double *coeff4, int32_t *kind4,
                                      int32_t *last4, int32_t *type2, int32_t *type3, int32_t *type4) {
    HYMD_REQUIRE(b, "NULL builder");
    HYMD_REQUIRE(b->total[0] == 0 || (a2 && b2 && eq2 && k2), "bond outputs are required");
    HYMD_REQUIRE(b->total[1] == 0 || (a3 && b3 && c3 && eq3 && k3), "angle outputs are required");
    HYMD_REQUIRE(b->total[2] == 0 || (a4 && b4 && c4 && d4 && coeff4 && kind4 && last4),
                 "dihedral outputs are required");
    if (b->n_mol == 0) return HYMD_OK;
    const int64_t nm = b->n_mol;
    int *off = (int *)b->offsets.ptr;
    TopoOut o{};
    o.off2 = off;
    o.off3 = off + (nm + 2);
    o.off4 = off + 2 * (nm + 2);
    o.a2 = a2, o.b2 = b2, o.eq2 = eq2, o.k2 = k2, o.t2 = type2;
    o.a3 = a3, o.b3 = b3, o.c3 = c3, o.eq3 = eq3, o.k3 = k3, o.t3 = type3;
    o.a4 = a4, o.b4 = b4, o.c4 = c4, o.d4 = d4, o.coef4 = coeff4, o.kind4 = kind4, o.last4 = last4, o.t4 = type4;
    topo_walk_kernel<true><<<grid_for(nm), kBlock, 0, b->stream>>>(b->g, b->t, o);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}

// ---------------------------------------------------------------------------
// prepare_index_based_bonds (force.py:508-570): every molecule copies the
// term list of its residue template with the indices offset by its first
// row.  One thread per molecule: counts, exclusive scans, fill.
// ---------------------------------------------------------------------------
namespace hymd {

struct IndexTemplates {
    int n_tmpl;
    const int *off2, *off3, *off4;  // [n_tmpl + 1] term ranges per template
    const int *i2, *i3, *i4;        // 0-based indices within the molecule, 2 / 3 / 4 per term
    const double *p2, *p3;          // (r0, k), (theta0 rad, k)
    const double *c4;               // [30] per dihedral
    const int *k4;                  // dih_type
};

__global__ void __launch_bounds__(kBlock) index_count_kernel(int64_t n_mol, const int *__restrict__ tmpl_of,
                                                             IndexTemplates t, int *__restrict__ cnt2,
                                                             int *__restrict__ cnt3, int *__restrict__ cnt4) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n_mol; m += stride) {
        const int r = tmpl_of[m];
        const bool ok = r >= 0 && r < t.n_tmpl;
        cnt2[m] = ok ? t.off2[r + 1] - t.off2[r] : 0;
        cnt3[m] = ok ? t.off3[r + 1] - t.off3[r] : 0;
        cnt4[m] = ok ? t.off4[r + 1] - t.off4[r] : 0;
    }
}

__global__ void __launch_bounds__(kBlock) index_fill_kernel(int64_t n_mol, const int64_t *__restrict__ first_row,
                                                            const int *__restrict__ tmpl_of, IndexTemplates t,
                                                            const int *__restrict__ o2, const int *__restrict__ o3,
                                                            const int *__restrict__ o4, TopoOut o) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n_mol; m += stride) {
        const int r = tmpl_of[m];
        if (r < 0 || r >= t.n_tmpl) continue;
        const int f = (int)first_row[m];
        for (int q = t.off2[r], w = o2[m]; q < t.off2[r + 1]; ++q, ++w) {
            o.a2[w] = f + t.i2[2 * q];
            o.b2[w] = f + t.i2[2 * q + 1];
            o.eq2[w] = t.p2[2 * q];
            o.k2[w] = t.p2[2 * q + 1];
        }
        for (int q = t.off3[r], w = o3[m]; q < t.off3[r + 1]; ++q, ++w) {
            o.a3[w] = f + t.i3[3 * q];
            o.b3[w] = f + t.i3[3 * q + 1];
            o.c3[w] = f + t.i3[3 * q + 2];
            o.eq3[w] = t.p3[2 * q];
            o.k3[w] = t.p3[2 * q + 1];
        }
        for (int q = t.off4[r], w = o4[m]; q < t.off4[r + 1]; ++q, ++w) {
            o.a4[w] = f + t.i4[4 * q];
            o.b4[w] = f + t.i4[4 * q + 1];
            o.c4[w] = f + t.i4[4 * q + 2];
            o.d4[w] = f + t.i4[4 * q + 3];
            o.kind4[w] = t.k4[q];
            o.last4[w] = 0;
            for (int c = 0; c < 30; ++c) o.coef4[(int64_t)w * 30 + c] = t.c4[30 * q + c];
        }
    }
}

}  // namespace hymd

struct hymd_index_bonds {
    hipStream_t stream = nullptr;
    int64_t n_mol = 0;
    const int64_t *first_row = nullptr;
    const int32_t *tmpl_of = nullptr;
    IndexTemplates t{};
    DevBuf tables, counts, offsets, tiles;
    int64_t total[3] = {0, 0, 0};
};

extern "C" int hymd_index_bonds_destroy(hymd_index_bonds *b) {
    if (!b) return HYMD_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
    return HYMD_OK;
}

extern "C" int hymd_index_bonds_create(hymd_index_bonds **out, int64_t n_mol, const int64_t *first_row,
                                       const int32_t *template_of, int32_t n_templates, const int32_t *offsets,
                                       const int32_t *idx2, const double *par2, const int32_t *idx3,
                                       const double *par3, const int32_t *idx4, const double *coeff4,
                                       const int32_t *kind4, int64_t counts[3], void *stream) {
    HYMD_REQUIRE(out && counts, "NULL output");
    *out = nullptr;
    HYMD_REQUIRE(n_mol >= 0 && n_templates >= 0 && offsets, "bad sizes");
    HYMD_REQUIRE(n_mol == 0 || (first_row && template_of), "NULL molecule arrays");
    const int nt = n_templates;
    const int n2 = offsets[nt], n3 = offsets[2 * (nt + 1) - 1], n4 = offsets[3 * (nt + 1) - 1];
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < nt; ++r)
            HYMD_REQUIRE(offsets[k * (nt + 1) + r] <= offsets[k * (nt + 1) + r + 1] && offsets[k * (nt + 1)] == 0,
                         "template offsets must start at 0 and not decrease");
    HYMD_REQUIRE((n2 == 0 || (idx2 && par2)) && (n3 == 0 || (idx3 && par3)) && (n4 == 0 || (idx4 && coeff4 && kind4)),
                 "NULL template table");
    auto *b = new hymd_index_bonds();
    b->stream = (hipStream_t)stream;
    b->n_mol = n_mol;
    b->first_row = first_row;
    b->tmpl_of = template_of;
    // one staging copy of the small template tables
    const size_t nd = (size_t)2 * n2 + 2 * n3 + 30 * n4 + 1;
    const size_t ni = (size_t)3 * (nt + 1) + 2 * n2 + 3 * n3 + 4 * n4 + n4 + 1;
    int st = b->tables.reserve(sizeof(double) * nd + sizeof(int) * ni);
    if (st == HYMD_OK) st = b->counts.reserve(sizeof(int) * 3 * (size_t)(n_mol + 1));
    if (st == HYMD_OK) st = b->offsets.reserve(sizeof(int) * 3 * (size_t)(n_mol + 2));
    if (st == HYMD_OK) st = b->tiles.reserve(sizeof(int) * (size_t)(scan_tiles(n_mol) + 1));
    if (st != HYMD_OK) {
        delete b;
        return st;
    }
    std::vector<double> hd;
    std::vector<int> hi;
    hd.insert(hd.end(), par2, par2 + 2 * n2);
    hd.insert(hd.end(), par3, par3 + 2 * n3);
    hd.insert(hd.end(), coeff4, coeff4 + 30 * n4);
    hd.push_back(0.0);
    hi.insert(hi.end(), offsets, offsets + 3 * (nt + 1));
    hi.insert(hi.end(), idx2, idx2 + 2 * n2);
    hi.insert(hi.end(), idx3, idx3 + 3 * n3);
    hi.insert(hi.end(), idx4, idx4 + 4 * n4);
    hi.insert(hi.end(), kind4, kind4 + n4);
    hi.push_back(0);
    double *td = (double *)b->tables.ptr;
    int *ti = (int *)(td + hd.size());
    hipStream_t s = b->stream;
    HYMD_HIP(hipMemcpyAsync(td, hd.data(), sizeof(double) * hd.size(), hipMemcpyHostToDevice, s));
    HYMD_HIP(hipMemcpyAsync(ti, hi.data(), sizeof(int) * hi.size(), hipMemcpyHostToDevice, s));
    IndexTemplates &t = b->t;
    t.n_tmpl = nt;
    t.off2 = ti;
    t.off3 = ti + (nt + 1);
    t.off4 = ti + 2 * (nt + 1);
    t.i2 = ti + 3 * (nt + 1);
    t.i3 = t.i2 + 2 * n2;
    t.i4 = t.i3 + 3 * n3;
    t.k4 = t.i4 + 4 * n4;
    t.p2 = td;
    t.p3 = td + 2 * n2;
    t.c4 = td + 2 * n2 + 2 * n3;
    int tot[3] = {0, 0, 0};
    if (n_mol > 0) {
        int *cnt = (int *)b->counts.ptr;
        int *off = (int *)b->offsets.ptr;
        index_count_kernel<<<grid_for(n_mol), kBlock, 0, s>>>(n_mol, template_of, t, cnt, cnt + (n_mol + 1),
                                                              cnt + 2 * (n_mol + 1));
        HYMD_HIP(hipGetLastError());
        for (int k = 0; k < 3; ++k)
            HYMD_TRY(exclusive_scan(cnt + k * (n_mol + 1), off + k * (n_mol + 2), n_mol, (int *)b->tiles.ptr,
                                    nullptr, false, s));
        for (int k = 0; k < 3; ++k)
            HYMD_HIP(hipMemcpyAsync(&tot[k], off + k * (n_mol + 2) + n_mol, sizeof(int), hipMemcpyDeviceToHost, s));
    }
    HYMD_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 3; ++k) b->total[k] = counts[k] = tot[k];
    *out = b;
    return HYMD_OK;
}

extern "C" int hymd_index_bonds_fill(hymd_index_bonds *b, int32_t *a2, int32_t *b2, double *eq2, double *k2,
                                     int32_t *a3, int32_t *b3, int32_t *c3, double *eq3, double *k3, int32_t *a4,
                                     int32_t *b4, int32_t *c4, int32_t *d4, double *coeff4, int32_t *kind4,
                                     int32_t *last4) {
    HYMD_REQUIRE(b, "NULL handle");
    HYMD_REQUIRE(b->total[0] == 0 || (a2 && b2 && eq2 && k2), "bond outputs are required");
    HYMD_REQUIRE(b->total[1] == 0 || (a3 && b3 && c3 && eq3 && k3), "angle outputs are required");
    HYMD_REQUIRE(b->total[2] == 0 || (a4 && b4 && c4 && d4 && coeff4 && kind4 && last4),
                 "dihedral outputs are required");
    if (b->n_mol == 0) return HYMD_OK;
    const int64_t nm = b->n_mol;
    const int *off = (const int *)b->offsets.ptr;
    TopoOut o{};
    o.a2 = a2, o.b2 = b2, o.eq2 = eq2, o.k2 = k2;
    o.a3 = a3, o.b3 = b3, o.c3 = c3, o.eq3 = eq3, o.k3 = k3;
    o.a4 = a4, o.b4 = b4, o.c4 = c4, o.d4 = d4, o.coef4 = coeff4, o.kind4 = kind4, o.last4 = last4;
    index_fill_kernel<<<grid_for(nm), kBlock, 0, b->stream>>>(nm, b->first_row, b->tmpl_of, b->t, off,
                                                              off + (nm + 2), off + 2 * (nm + 2), o);
    HYMD_HIP(hipGetLastError());
    return HYMD_OK;
}
