// Bonded-term math shared by the bond / angle / dihedral kernels (device, FP64).
//
// Restates compute_bond_forces.f90:45-61, compute_angle_forces.f90:54-93,
// compute_dihedral_forces.f90:29-136 and dipole_reconstruction.f90:37-220 as
// per-term device functions.  The kernels gather: each particle walks its own
// list of (term, role) entries in increasing term order and recomputes the
// terms it belongs to, so forces need no atomics and match the reference's
// per-particle summation order.
#pragma once
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

__device__ __forceinline__ F3 bond_term_f(F3 rab, float r0, float k) {
    const float rn = sqrtf(dotf(rab, rab));
    const float s = -(k * (rn - r0)) / rn;
    return f3(s * rab.x, s * rab.y, s * rab.z);
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
    const float nn = sqrtf(dotf(n, n));
    if (!(nn > 0.0f)) {
        fa = fc = f3(0.f, 0.f, 0.f);
        return;
    }
    const float theta = atan2f(nn, dotf(ra, rc));
    const float g = -k * (theta - t0) / nn;
    const float ga = g / dotf(ra, ra), gc = g / dotf(rc, rc);
    fa = f3(ga * (n.y * ra.z - n.z * ra.y), ga * (n.z * ra.x - n.x * ra.z), ga * (n.x * ra.y - n.y * ra.x));
    fc = f3(gc * (rc.y * n.z - rc.z * n.y), gc * (rc.z * n.x - rc.x * n.z), gc * (rc.x * n.y - rc.y * n.x));
}

// --- cosine series  E += sum c_n (1 + cos(n phi - d_n)), dE/dphi ----------
__device__ __forceinline__ void cosine_series(const float *c, const float *d, double phi, double &e, double &de) {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        const double cn = (double)c[i], dn = (double)d[i];
        e = e + cn * (1.0 + cos(i * phi - dn));
        de = de - i * cn * sin(i * phi - dn);
    }
}

struct Mat3 {
    double m[3][3];
};
__device__ __forceinline__ Mat3 cross_matrix(const Mat3 &M, V3 v) {
    Mat3 o;
    for (int i = 0; i < 3; ++i) {
        V3 row = v3(M.m[i][0], M.m[i][1], M.m[i][2]);
        V3 c = cross(row, v);
        o.m[i][0] = c.x;
        o.m[i][1] = c.y;
        o.m[i][2] = c.z;
    }
    return o;
}
__device__ __forceinline__ Mat3 outer(V3 a, V3 b) {
    Mat3 o;
    for (int i = 0; i < 3; ++i) {
        const double ai = get(a, i);
        o.m[i][0] = ai * b.x;
        o.m[i][1] = ai * b.y;
        o.m[i][2] = ai * b.z;
    }
    return o;
}

struct Recon {
    double energy, df;
    V3 fa, fb, fc;
    V3 dip[2];
    double tm[3][3][3];
    bool clamped;
    bool has_dipole;
};

// dipole_reconstruction.f90:50-220 (combined bending-torsion + dipoles).
// Outputs are zero where the reference leaves them undefined (cos^2 >= 1).
__device__ __noinline__ void reconstruct(V3 rab, V3 rb, V3 rcb, const Box &box, const float *ck, const float *dk_,
                                         double phi, int dipole_flag, Recon &o) {
    const double delta = 0.3;
    const double cos_psi = (double)cosf(1.392947f), sin_psi = (double)sinf(1.392947f);
    const double gamma_0 = 1.85 - 0.227 * cos(phi - 0.785);
    const double dg = 0.227 * sin(phi - 0.785);
    double k = 0.0, dk = 0.0;
    cosine_series(ck, dk_, phi, k, dk);
    const double norm_a = norm(rab), norm_c = norm(rcb);
    const V3 w = rab / norm_a, v = rcb / norm_c;
    const double cos_gamma = dot(w, v);
    const double cos2 = cos_gamma * cos_gamma;
    o.energy = 0.0;
    o.df = 0.0;
    o.fa = o.fb = o.fc = v3(0, 0, 0);
    o.clamped = false;
    o.has_dipole = false;
    if (!(cos2 < 1.0)) return;
    const double gamm = acos(cos_gamma);
    double sin_gamma = sqrt(1 - cos2);
    if (sin_gamma < 0.1) {
        o.clamped = true;
        sin_gamma = (double)0.1f;
    }
    V3 fa = -((v - cos_gamma * w) / norm_a) / sin_gamma;
    V3 fc = -((w - cos_gamma * v) / norm_c) / sin_gamma;
    V3 fb = -(fa + fc);
    const double df_ang = k * (gamm - gamma_0);
    const double var_sq = (gamm - gamma_0) * (gamm - gamma_0);
    o.energy = 0.5 * k * var_sq;
    o.df = 0.5 * dk * var_sq - df_ang * dg;
    if (dipole_flag != 0) {
        o.has_dipole = true;
        const double fac = exp((gamm - 1.73) / 0.025);
        const double theta = -1.607 * gamm + 0.094 + 1.883 / (1.0 + fac);
        const double d_theta = -1.607 - 1.883 / 0.025 * fac / ((1.0 + fac) * (1.0 + fac));
        const double cos_theta = cos(theta), sin_theta = sin(theta);
        const V3 n = cross(w, v) / sin_gamma;
        const V3 m = cross(n, v);
        const V3 r0 = rb + 0.5 * rcb;
        const V3 d = (0.5 * delta) * (cos_psi * v + sin_psi * (cos_theta * n + sin_theta * m));
        V3 d1 = r0 + d, d2 = r0 - d;
        o.dip[0] = mic(d1, box);
        o.dip[1] = mic(d2, box);
        Mat3 Vb, Wb, Vc, Wa;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double vb = get(v, i) * get(v, j), wb = get(w, i) * get(w, j);
                if (i == j) {
                    vb -= 1.0;
                    wb -= 1.0;
                }
                Vb.m[i][j] = vb / norm_c;
                Wb.m[i][j] = wb / norm_a;
                Vc.m[i][j] = -Vb.m[i][j];
                Wa.m[i][j] = -Wb.m[i][j];
            }
        const Mat3 oa = outer(fa, n), ob = outer(fb, n), oc = outer(fc, n);
        const Mat3 cWa = cross_matrix(Wa, v), cWb = cross_matrix(Wb, v), cVb = cross_matrix(Vb, w),
                   cVc = cross_matrix(Vc, w);
        Mat3 Na, Nb, Nc;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                Na.m[i][j] = (cos_gamma * oa.m[i][j] + cWa.m[i][j]) / sin_gamma;
                Nb.m[i][j] = (cos_gamma * ob.m[i][j] + cWb.m[i][j] - cVb.m[i][j]) / sin_gamma;
                Nc.m[i][j] = (cos_gamma * oc.m[i][j] - cVc.m[i][j]) / sin_gamma;
            }
        const Mat3 Ma = cross_matrix(Na, v);
        const Mat3 t1 = cross_matrix(Nb, v), t2 = cross_matrix(Vb, n);
        const Mat3 t3 = cross_matrix(Nc, v), t4 = cross_matrix(Vc, n);
        const Mat3 qa = outer(fa, m), qb = outer(fb, m), qc = outer(fc, m);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const double Mb = t1.m[i][j] - t2.m[i][j], Mc = t3.m[i][j] - t4.m[i][j];
                const double FNa = sin_theta * d_theta * oa.m[i][j], FNb = sin_theta * d_theta * ob.m[i][j],
                             FNc = sin_theta * d_theta * oc.m[i][j];
                const double FMa = cos_theta * d_theta * qa.m[i][j], FMb = cos_theta * d_theta * qb.m[i][j],
                             FMc = cos_theta * d_theta * qc.m[i][j];
                o.tm[0][i][j] =
                    0.5 * delta * (sin_psi * (cos_theta * Na.m[i][j] + sin_theta * Ma.m[i][j] + FNa - FMa));
                o.tm[1][i][j] = 0.5 * delta *
                                (cos_psi * Vb.m[i][j] + sin_psi * (cos_theta * Nb.m[i][j] + sin_theta * Mb + FNb - FMb));
                o.tm[2][i][j] = 0.5 * delta *
                                (cos_psi * Vc.m[i][j] + sin_psi * (cos_theta * Nc.m[i][j] + sin_theta * Mc + FNc - FMc));
            }
    }
    o.fa = df_ang * fa;
    o.fb = df_ang * fb;
    o.fc = df_ang * fc;
}

}  // namespace hymd
