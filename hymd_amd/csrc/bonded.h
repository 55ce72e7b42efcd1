// Bonded-term math shared by the bond / angle / dihedral kernels (device, FP64).
//
// Restates compute_bond_forces.f90:45-61, compute_angle_forces.f90:54-93,
// compute_dihedral_forces.f90:29-136 and dipole_reconstruction.f90:37-220 as
// per-term device functions.  The kernels gather: each particle walks its own
// list of (term, role) entries in increasing term order and recomputes the
// terms it belongs to, so forces need no atomics and match the reference's
// per-particle summation order.
#pragma once
#include <cmath>

#include "common.h"

namespace hymd {

struct V3 {
    double x, y, z;
};
__device__ __forceinline__ V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 operator/(V3 a, double s) { return v3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double norm(V3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ double get(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

struct Box {
    double l[3];
    double inv[3];  // 1 / l
};
inline Box make_box3(const double *l) {
    Box b;
    for (int d = 0; d < 3; ++d) {
        b.l[d] = l[d];
        b.inv[d] = 1.0 / l[d];
    }
    return b;
}

// minimum image, Fortran NINT (round half away from zero) of d / L, with the
// quotient taken as d * (1/L)
__device__ __forceinline__ V3 mic(V3 v, const Box &b) {
    return v3(v.x - b.l[0] * round(v.x * b.inv[0]), v.y - b.l[1] * round(v.y * b.inv[1]),
              v.z - b.l[2] * round(v.z * b.inv[2]));
}

template <typename PosT>
__device__ __forceinline__ V3 load3(const PosT *__restrict__ r, int64_t i) {
    return v3((double)r[3 * i], (double)r[3 * i + 1], (double)r[3 * i + 2]);
}

// --- harmonic bond: returns fa (force on b; a gets -fa) ------------------
__device__ __forceinline__ V3 bond_term(V3 rab, double r0, double k, double &energy) {
    const double rn = norm(rab);
    const double dr = rn - r0;
    const double df = k * dr;
    energy = 0.5 * k * (dr * dr);
    const double s = -df / rn;
    return v3(s * rab.x, s * rab.y, s * rab.z);
}

// --- harmonic angle in theta: fa, fc (a gets -fa, c -fc, b +fa+fc) --------
// sin(theta) = sqrt((1 - cos)(1 + cos)) (both factors exact near cos = +-1),
// and two divisions (1 / (|ra||rc|), ff / sin) instead of eight.
__device__ __forceinline__ bool angle_term(V3 ra, V3 rc, double t0, double k, V3 &fa, V3 &fc, double &energy) {
    const double na = norm(ra), nc = norm(rc);
    const double inv = 1.0 / (na * nc);
    const double ina = nc * inv, inc = na * inv;
    const V3 ea = ina * ra, ec = inc * rc;
    const double cosphi = dot(ea, ec);
    if (!(cosphi * cosphi < 1.0)) {
        fa = fc = v3(0, 0, 0);
        energy = 0.0;
        return false;
    }
    const double theta = acos(cosphi);
    const double sinphi = sqrt((1.0 - cosphi) * (1.0 + cosphi));
    const double d = theta - t0;
    const double ff = k * d;
    const double q = -ff / sinphi;
    fa = (q * ina) * (ec - cosphi * ea);
    fc = (q * inc) * (ea - cosphi * ec);
    energy = 0.5 * ff * d;
    return true;
}

// --- the same two terms in FP32 arithmetic (this path's FP32 forces; the
// reference's real(4) kernels compute_bond_forces.f90 / compute_angle_forces.f90
// difference positions in FP32 and evaluate the terms in FP64), on
// separations taken in FP64 and rounded once; no energies (the energy pass
// stays FP64) ------
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 to_f3(V3 a) { return F3{(float)a.x, (float)a.y, (float)a.z}; }
__device__ __forceinline__ float dotf(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// FP32 term arithmetic takes the hardware reciprocal square root and
// reciprocal (1 ulp each) instead of the correctly rounded square root and
// division sequences (about ten instructions each): the terms stay within a
// few FP32 ulps, and the fused inner loop, bound by VALU issue, runs fewer
// instructions per sub-step.
__device__ __forceinline__ F3 bond_term_f(F3 rab, float r0, float k) {
    const float r2 = dotf(rab, rab);
    const float ri = __builtin_amdgcn_rsqf(r2);
    const float s = -(k * (r2 * ri - r0)) * ri;
    return f3(s * rab.x, s * rab.y, s * rab.z);
}

// atan2(y, x) for y > 0 in FP32: t = min(|x|, y) / max(|x|, y) in [0, 1]
// (hardware reciprocal and one Newton step), atan(t) = t P(t^2) with P of
// degree 8 fitted to atan(sqrt s) / sqrt s on s in [0, 1] (minimax in the
// relative error by Lawson's iteratively reweighted least squares on 4000
// Chebyshev-spaced points; 1.6e-8 relative in exact arithmetic), then the
// octant fixes pi/2 - a (y > |x|) and pi - a (x < 0).  About 22 VALU
// instructions against 45 for atan2f, within a few FP32 ulps of atan2.
constexpr float kAtanC[9] = {0x1p+0f,        -0x1.5554a6p-2f, 0x1.9972e8p-3f, -0x1.22e402p-3f, 0x1.b3da48p-4f,
                             -0x1.33603cp-4f, 0x1.5dba9ap-5f,  -0x1.0744b2p-6f, 0x1.758a5ep-9f};

__device__ __forceinline__ float atan2_ypos(float y, float x) {
    const float ax = fabsf(x);
    const float mx = fmaxf(ax, y), mn = fminf(ax, y);
    float r = __builtin_amdgcn_rcpf(mx);
    r = r * fmaf(-mx, r, 2.0f);
    const float t = mn * r, s = t * t;
    float p = kAtanC[8];
#pragma unroll
    for (int k = 7; k >= 0; --k) p = fmaf(p, s, kAtanC[k]);
    float a = t * p;
    if (y > ax) a = 0x1.921fb6p+0f - a;  // pi / 2
    if (x < 0.0f) a = 0x1.921fb6p+1f - a;  // pi
    return a;
}

// theta = atan2(|n|, ra.rc) with n = ra x rc formed in FP64 from the FP64
// separations, and the force directions n x ra and rc x n (perpendicular to
// ra and rc, so free of cancellation): the error of theta and of the forces
// stays at FP32 rounding for nearly straight angles too, where acos(cos)
// would amplify it by 1 / sin(theta).
//   fa = -k (theta - t0) (n x ra) / (|n| |ra|^2),  fc = -k (theta - t0) (rc x n) / (|n| |rc|^2)
__device__ __forceinline__ void angle_term_f(V3 ra64, V3 rc64, float t0, float k, F3 &fa, F3 &fc) {
    const F3 n = to_f3(cross(ra64, rc64));
    const F3 ra = to_f3(ra64), rc = to_f3(rc64);
    const float n2 = dotf(n, n);
    const float ni = __builtin_amdgcn_rsqf(n2);
    const float nn = n2 * ni;  // NaN for n = 0 (ni = inf): no force, as below
    if (!(nn > 0.0f)) {
        fa = fc = f3(0.f, 0.f, 0.f);
        return;
    }
    const float theta = atan2_ypos(nn, dotf(ra, rc));
    const float g = -k * (theta - t0) * ni;
    const float ga = g * __builtin_amdgcn_rcpf(dotf(ra, ra)), gc = g * __builtin_amdgcn_rcpf(dotf(rc, rc));
    fa = f3(ga * (n.y * ra.z - n.z * ra.y), ga * (n.z * ra.x - n.x * ra.z), ga * (n.x * ra.y - n.y * ra.x));
    fc = f3(gc * (rc.y * n.z - rc.z * n.y), gc * (rc.z * n.x - rc.x * n.z), gc * (rc.x * n.y - rc.y * n.x));
}

// --- cosine series  E += sum c_n (1 + cos(n phi - d_n)), dE/dphi ----------
// (one argument reduction per term: sincos; the loop kept rolled so that one
// copy of the FP64 sin/cos sequence holds registers)
__device__ __forceinline__ void cosine_series(const float *c, const float *d, double phi, double &e, double &de) {
#pragma unroll 1
    for (int i = 0; i < 5; ++i) {
        const double cn = (double)c[i], dn = (double)d[i];
        double sn, cs;
        sincos(i * phi - dn, &sn, &cs);
        e = e + cn * (1.0 + cs);
        de = de - i * cn * sn;
    }
}

// Geometry of one combined bending-torsion triple (dipole_reconstruction.f90
// :62-107): gamma between r_ab and r_cb, the unit force directions of gamma
// (fa on a, fc on c, fb = -(fa + fc)), sin(gamma) clamped at 0.1.  Returns
// false where cos^2(gamma) >= 1 (the reference leaves its outputs undefined
// there; they are zero here).
struct CbtGeom {
    double norm_a, norm_c, cos_gamma, sin_gamma, gamm;
    V3 w, v, fa, fb, fc;
    bool clamped;
};
__device__ __forceinline__ bool cbt_geometry(V3 rab, V3 rcb, CbtGeom &o) {
    o.norm_a = norm(rab);
    o.norm_c = norm(rcb);
    o.w = rab / o.norm_a;
    o.v = rcb / o.norm_c;
    o.cos_gamma = dot(o.w, o.v);
    const double cos2 = o.cos_gamma * o.cos_gamma;
    if (!(cos2 < 1.0)) return false;
    o.gamm = acos(o.cos_gamma);
    o.sin_gamma = sqrt(1 - cos2);
    o.clamped = o.sin_gamma < 0.1;
    if (o.clamped) o.sin_gamma = (double)0.1f;
    o.fa = -((o.v - o.cos_gamma * o.w) / o.norm_a) / o.sin_gamma;
    o.fc = -((o.w - o.cos_gamma * o.v) / o.norm_c) / o.sin_gamma;
    o.fb = -(o.fa + o.fc);
    return true;
}

// The dipole pair and transfer matrices of one combined bending-torsion
// triple (dipole_reconstruction.f90:148-213), written to dip6[6] and
// tm27[27] (matrices for a, b, c, row-major) row by row: the 3x3 products
// of the reference (outer products, row-wise cross products) are formed one
// row at a time, entry by entry in the reference's arithmetic, so no matrix
// is held whole.  w, v: unit vectors along r_ab and r_cb; fa, fb, fc: the
// unit force directions of gamma; rb: the middle bead.
template <typename OutT>
__device__ __forceinline__ void cbt_dipole_outputs(V3 w, V3 v, V3 fa, V3 fb, V3 fc, double cos_gamma,
                                                   double sin_gamma, double gamm, double norm_a, double norm_c, V3 rb,
                                                   V3 rcb, const Box &box, OutT *__restrict__ dip6,
                                                   OutT *__restrict__ tm27) {
    const double delta = 0.3;
    const double cos_psi = (double)cosf(1.392947f), sin_psi = (double)sinf(1.392947f);
    const double fac = exp((gamm - 1.73) / 0.025);
    const double theta = -1.607 * gamm + 0.094 + 1.883 / (1.0 + fac);
    const double d_theta = -1.607 - 1.883 / 0.025 * fac / ((1.0 + fac) * (1.0 + fac));
    const double cos_theta = cos(theta), sin_theta = sin(theta);
    const V3 n = cross(w, v) / sin_gamma;
    const V3 m = cross(n, v);
    const V3 r0 = rb + 0.5 * rcb;
    const V3 d = (0.5 * delta) * (cos_psi * v + sin_psi * (cos_theta * n + sin_theta * m));
    const V3 d1 = mic(r0 + d, box), d2 = mic(r0 - d, box);
    dip6[0] = (OutT)d1.x;
    dip6[1] = (OutT)d1.y;
    dip6[2] = (OutT)d1.z;
    dip6[3] = (OutT)d2.x;
    dip6[4] = (OutT)d2.y;
    dip6[5] = (OutT)d2.z;
#pragma unroll 1
    for (int i = 0; i < 3; ++i) {  // one row's state live at a time
        // row i of V_b = (v v^T - 1) / |r_cb|, W_b = (w w^T - 1) / |r_ab|, V_c = -V_b, W_a = -W_b
        double vb[3], wb[3];
        for (int j = 0; j < 3; ++j) {
            double x = get(v, i) * get(v, j), y = get(w, i) * get(w, j);
            if (i == j) {
                x -= 1.0;
                y -= 1.0;
            }
            vb[j] = x / norm_c;
            wb[j] = y / norm_a;
        }
        const V3 Vb = v3(vb[0], vb[1], vb[2]), Wb = v3(wb[0], wb[1], wb[2]);
        const V3 Vc = v3(-vb[0], -vb[1], -vb[2]), Wa = v3(-wb[0], -wb[1], -wb[2]);
        // rows i of fa n^T, fb n^T, fc n^T and of the cross-product matrices
        const double fai = get(fa, i), fbi = get(fb, i), fci = get(fc, i);
        const V3 oa = fai * n, ob = fbi * n, oc = fci * n;
        const V3 cWa = cross(Wa, v), cWb = cross(Wb, v), cVb = cross(Vb, w), cVc = cross(Vc, w);
        double na[3], nb[3], nc[3];
        for (int j = 0; j < 3; ++j) {
            na[j] = (cos_gamma * get(oa, j) + get(cWa, j)) / sin_gamma;
            nb[j] = (cos_gamma * get(ob, j) + get(cWb, j) - get(cVb, j)) / sin_gamma;
            nc[j] = (cos_gamma * get(oc, j) - get(cVc, j)) / sin_gamma;
        }
        const V3 Na = v3(na[0], na[1], na[2]), Nb = v3(nb[0], nb[1], nb[2]), Nc = v3(nc[0], nc[1], nc[2]);
        const V3 Ma = cross(Na, v);
        const V3 t1 = cross(Nb, v), t2 = cross(Vb, n);
        const V3 t3 = cross(Nc, v), t4 = cross(Vc, n);
        const V3 qa = fai * m, qb = fbi * m, qc = fci * m;
        for (int j = 0; j < 3; ++j) {
            const double Mb = get(t1, j) - get(t2, j), Mc = get(t3, j) - get(t4, j);
            const double FNa = sin_theta * d_theta * get(oa, j), FNb = sin_theta * d_theta * get(ob, j),
                         FNc = sin_theta * d_theta * get(oc, j);
            const double FMa = cos_theta * d_theta * get(qa, j), FMb = cos_theta * d_theta * get(qb, j),
                         FMc = cos_theta * d_theta * get(qc, j);
            tm27[3 * i + j] = (OutT)(0.5 * delta * (sin_psi * (cos_theta * na[j] + sin_theta * get(Ma, j) + FNa - FMa)));
            tm27[9 + 3 * i + j] =
                (OutT)(0.5 * delta * (cos_psi * vb[j] + sin_psi * (cos_theta * nb[j] + sin_theta * Mb + FNb - FMb)));
            tm27[18 + 3 * i + j] =
                (OutT)(0.5 * delta * (cos_psi * get(Vc, j) + sin_psi * (cos_theta * nc[j] + sin_theta * Mc + FNc - FMc)));
        }
    }
}

// One dihedral term (compute_dihedral_forces.f90:35-134) on the positions of
// its four particles: the force on each (roles 0..3 = a, b, c, d), its energy
// and clamp count into ev[0], ev[1], and its dipoles / transfer matrices into
// dip12[12] / tm54[54] (zero where the term has none, as the reference's
// zeroed outputs); cf = the term's 30 coefficients, dt its type, last its
// second-reconstruction flag.  dip12 == nullptr (with dipole_flag 0) writes
// no dipoles.
template <typename OutT>
__device__ void dihedral_core(V3 ra, V3 rb, V3 rc, V3 rd, const float *__restrict__ cf, int dt, int last,
                              const Box &box, int dipole_flag, double *ev, OutT *__restrict__ dip12,
                              OutT *__restrict__ tm54, V3 (&out)[4]) {
    const V3 f = mic(ra - rb, box), g = mic(rb - rc, box), h = mic(rd - rc, box);
    const V3 v = cross(f, g), w = cross(h, g);
    const double v_sq = dot(v, v), w_sq = dot(w, w), g_norm = norm(g);
    const double cos_phi = dot(v, w);
    const double sin_phi = dot(w, f) * g_norm;
    const double phi = atan2(sin_phi, cos_phi);
    const double f_dot_g = dot(f, g), h_dot_g = dot(h, g);
    double energy = 0.0, df = 0.0;
    for (int q = 0; q < 4; ++q) out[q] = v3(0, 0, 0);
    if (dt == 0 || dt == 1) {
        cosine_series(cf, cf + 5, phi, energy, df);
        int zc = 0, zd = 0;
        for (int i = 0; i < 5; ++i) {
            zc += (cf[10 + i] == 0.f);
            zd += (cf[15 + i] == 0.f);
        }
        if (zc != 5 && zd != 5) cosine_series(cf + 10, cf + 15, phi, energy, df);
    }
    // combined bending-torsion (dipole_reconstruction.f90:50-220) on the
    // triples (a, b, c) and, with last == 1, (b, c, d): K(phi) and
    // gamma_0(phi) are shared by both
    bool dip_done[2] = {false, false};
    if (dt == 1) {
        const double gamma_0 = 1.85 - 0.227 * cos(phi - 0.785);
        const double dg = 0.227 * sin(phi - 0.785);
        double kc = 0.0, dkc = 0.0;
        cosine_series(cf + 20, cf + 25, phi, kc, dkc);
        const int ntrip = last == 1 ? 2 : 1;
#pragma unroll 1
        for (int r = 0; r < ntrip; ++r) {
            const V3 rab = r == 0 ? f : g, rcb = r == 0 ? -g : h, rmid = r == 0 ? rb : rc;
            CbtGeom cg;
            if (!cbt_geometry(rab, rcb, cg)) continue;
            if (cg.clamped) ev[1] += 1.0;
            const double df_ang = kc * (cg.gamm - gamma_0);
            const double var_sq = (cg.gamm - gamma_0) * (cg.gamm - gamma_0);
            energy += 0.5 * kc * var_sq;
            df += 0.5 * dkc * var_sq - df_ang * dg;
            // constant indices (a register array indexed by r would live in scratch)
            if (r == 0) {
                out[0] = out[0] - df_ang * cg.fa;
                out[1] = out[1] - df_ang * cg.fb;
                out[2] = out[2] - df_ang * cg.fc;
            } else {
                out[1] = out[1] - df_ang * cg.fa;
                out[2] = out[2] - df_ang * cg.fb;
                out[3] = out[3] - df_ang * cg.fc;
            }
            if (dipole_flag != 0) {
                cbt_dipole_outputs(cg.w, cg.v, cg.fa, cg.fb, cg.fc, cg.cos_gamma, cg.sin_gamma, cg.gamm, cg.norm_a,
                                   cg.norm_c, rmid, rcb, box, dip12 + 6 * r, tm54 + 27 * r);
                dip_done[r] = true;
            }
        }
    }
    for (int r = 0; r < 2; ++r)
        if (dip12 && !dip_done[r]) {
            for (int k = 0; k < 6; ++k) dip12[6 * r + k] = (OutT)0;
            for (int k = 0; k < 27; ++k) tm54[27 * r + k] = (OutT)0;
        }
    if (dt == 2) {
        const double eq = (double)cf[0], kf = (double)cf[1];
        df = kf * (phi - eq);
        energy += 0.5 * kf * ((phi - eq) * (phi - eq));
    }
    const V3 sc = (f_dot_g / (v_sq * g_norm)) * v - (h_dot_g / (w_sq * g_norm)) * w;
    const V3 fa = (-df * g_norm / v_sq) * v;
    const V3 fd = (df * g_norm / w_sq) * w;
    out[0] = out[0] + fa;
    out[1] = out[1] + (df * sc - fa);
    out[2] = out[2] + (-df * sc - fd);
    out[3] = out[3] + fd;
    ev[0] += energy;
}


// The dipoles and transfer matrices of one dihedral term alone: the outputs
// of dihedral_core with dipole_flag 1, without its forces and energy.
template <typename OutT>
__device__ void dihedral_dipoles_core(V3 ra, V3 rb, V3 rc, V3 rd, int dt, int last, const Box &box,
                                      OutT *__restrict__ dip12, OutT *__restrict__ tm54) {
    const V3 f = mic(ra - rb, box), g = mic(rb - rc, box), h = mic(rd - rc, box);
    bool done[2] = {false, false};
    if (dt == 1) {
        const int ntrip = last == 1 ? 2 : 1;
#pragma unroll 1
        for (int r = 0; r < ntrip; ++r) {
            const V3 rab = r == 0 ? f : g, rcb = r == 0 ? -g : h, rmid = r == 0 ? rb : rc;
            CbtGeom cg;
            if (!cbt_geometry(rab, rcb, cg)) continue;
            cbt_dipole_outputs(cg.w, cg.v, cg.fa, cg.fb, cg.fc, cg.cos_gamma, cg.sin_gamma, cg.gamm, cg.norm_a,
                               cg.norm_c, rmid, rcb, box, dip12 + 6 * r, tm54 + 27 * r);
            done[r] = true;
        }
    }
    for (int r = 0; r < 2; ++r)
        if (!done[r]) {
            for (int k = 0; k < 6; ++k) dip12[6 * r + k] = (OutT)0;
            for (int k = 0; k < 27; ++k) tm54[27 * r + k] = (OutT)0;
        }
}


// The three cosine series of a dihedral term (the two torsion series and the
// stiffness K(phi) of the combined bending-torsion term), each
// sum_{n=0..4} c_n (1 + cos(n phi - d_n)), recast as
//   k0 + sum_{n=1..4} (a_n cos n phi + b_n sin n phi),
//   a_n = c_n cos d_n, b_n = c_n sin d_n, k0 = sum_n c_n + a_0,
// with d/dphi = sum_n n (b_n cos n phi - a_n sin n phi): the same function
// evaluated from cos phi and sin phi by the harmonic recurrence instead of two
// FP64 sine / cosine evaluations per harmonic.  Constants per term:
// kDihSeries doubles, series s at 9 s: (k0, a_1..a_4, b_1..b_4); a torsion
// series that dihedral_core skips (all c or all d zero) is stored as zeros.
constexpr int kDihSeries = 27;

// Host: the constants of one term from its 30 coefficients (c and d of the
// two torsion series and of K(phi): rows 0-1, 2-3, 4-5 of (6,5)).
inline void dihedral_series_constants(const float *cf, double *out27) {
    int zc = 0, zd = 0;
    for (int i = 0; i < 5; ++i) {
        zc += (cf[10 + i] == 0.f);
        zd += (cf[15 + i] == 0.f);
    }
    const bool second = zc != 5 && zd != 5;
    for (int sr = 0; sr < 3; ++sr) {
        double *o = out27 + 9 * sr;
        for (int k = 0; k < 9; ++k) o[k] = 0.0;
        if (sr == 1 && !second) continue;
        const float *c = cf + 10 * sr, *d = c + 5;
        double k0 = 0.0;
        for (int n = 0; n < 5; ++n) {
            const double a = (double)c[n] * std::cos((double)d[n]), b = (double)c[n] * std::sin((double)d[n]);
            k0 += (double)c[n];
            if (n == 0) k0 += a;
            else {
                o[n] = a;
                o[4 + n] = b;
            }
        }
        o[0] = k0;
    }
}

// One series (constants ser[0], ser[st], ... ser[8 st]) at cos n phi, sin n phi.
__device__ __forceinline__ void series_from_harmonics(const double *__restrict__ ser, int st, const double (&cn)[5],
                                                      const double (&sn)[5], double &e, double &de) {
    double ee = ser[0], dd = 0.0;
#pragma unroll
    for (int n = 1; n < 5; ++n) {
        const double a = ser[n * st], b = ser[(4 + n) * st];
        ee += a * cn[n] + b * sn[n];
        dd += n * (b * cn[n] - a * sn[n]);
    }
    e += ee;
    de += dd;
}

// dihedral_core's forces and energy (no dipoles) with the cosine series from
// their constants (ser, stride st between constants) and the harmonics of phi
// by recurrence; cf (the term's 30 coefficients) is read for dt == 2 only.
__device__ __forceinline__ void dihedral_core_series(V3 ra, V3 rb, V3 rc, V3 rd, const double *__restrict__ ser,
                                                     int st, const float *__restrict__ cf, int dt, int last,
                                                     const Box &box, double *ev, V3 (&out)[4]) {
    const V3 f = mic(ra - rb, box), g = mic(rb - rc, box), h = mic(rd - rc, box);
    const V3 v = cross(f, g), w = cross(h, g);
    const double v_sq = dot(v, v), w_sq = dot(w, w), g_norm = norm(g);
    const double cos_phi = dot(v, w);
    const double sin_phi = dot(w, f) * g_norm;
    const double f_dot_g = dot(f, g), h_dot_g = dot(h, g);
    double energy = 0.0, df = 0.0;
    for (int q = 0; q < 4; ++q) out[q] = v3(0, 0, 0);
    if (dt == 0 || dt == 1) {
        // cos phi, sin phi of phi = atan2(sin_phi, cos_phi) (phi = 0 when both vanish)
        const double rr = sqrt(cos_phi * cos_phi + sin_phi * sin_phi);
        double cn[5], sn[5];
        cn[0] = 1.0;
        sn[0] = 0.0;
        cn[1] = rr > 0.0 ? cos_phi / rr : 1.0;
        sn[1] = rr > 0.0 ? sin_phi / rr : 0.0;
#pragma unroll
        for (int n = 2; n < 5; ++n) {
            cn[n] = cn[n - 1] * cn[1] - sn[n - 1] * sn[1];
            sn[n] = sn[n - 1] * cn[1] + cn[n - 1] * sn[1];
        }
        series_from_harmonics(ser, st, cn, sn, energy, df);
        series_from_harmonics(ser + 9 * st, st, cn, sn, energy, df);
        if (dt == 1) {
            // gamma_0(phi) = 1.85 - 0.227 cos(phi - 0.785) and its derivative, by the angle sum
            const double c785 = 0x1.6a2ecb934b59ap-1, s785 = 0x1.69e4fd79ac743p-1;  // cos, sin of 0.785
            const double gamma_0 = 1.85 - 0.227 * (cn[1] * c785 + sn[1] * s785);
            const double dg = 0.227 * (sn[1] * c785 - cn[1] * s785);
            double kc = 0.0, dkc = 0.0;
            series_from_harmonics(ser + 18 * st, st, cn, sn, kc, dkc);
            const int ntrip = last == 1 ? 2 : 1;
#pragma unroll 1
            for (int r = 0; r < ntrip; ++r) {
                const V3 rab = r == 0 ? f : g, rcb = r == 0 ? -g : h;
                CbtGeom cg;
                if (!cbt_geometry(rab, rcb, cg)) continue;
                if (cg.clamped) ev[1] += 1.0;
                const double df_ang = kc * (cg.gamm - gamma_0);
                const double var_sq = (cg.gamm - gamma_0) * (cg.gamm - gamma_0);
                energy += 0.5 * kc * var_sq;
                df += 0.5 * dkc * var_sq - df_ang * dg;
                if (r == 0) {
                    out[0] = out[0] - df_ang * cg.fa;
                    out[1] = out[1] - df_ang * cg.fb;
                    out[2] = out[2] - df_ang * cg.fc;
                } else {
                    out[1] = out[1] - df_ang * cg.fa;
                    out[2] = out[2] - df_ang * cg.fb;
                    out[3] = out[3] - df_ang * cg.fc;
                }
            }
        }
    }
    if (dt == 2) {
        const double phi = atan2(sin_phi, cos_phi);
        const double eq = (double)cf[0], kf = (double)cf[1];
        df = kf * (phi - eq);
        energy += 0.5 * kf * ((phi - eq) * (phi - eq));
    }
    const V3 sc = (f_dot_g / (v_sq * g_norm)) * v - (h_dot_g / (w_sq * g_norm)) * w;
    const V3 fa = (-df * g_norm / v_sq) * v;
    const V3 fd = (df * g_norm / w_sq) * w;
    out[0] = out[0] + fa;
    out[1] = out[1] + (df * sc - fa);
    out[2] = out[2] + (-df * sc - fd);
    out[3] = out[3] + fd;
    ev[0] += energy;
}

}  // namespace hymd
