/*
 * C restatement of HyMD's Fortran bonded kernels -- ORACLE, TEST INFRASTRUCTURE ONLY.
 *
 * One serial loop per kernel, in term order, exactly as the reference:
 *   oracle_cbf  <- compute_bond_forces.f90:42-61      (cbf / cbf_d)
 *   oracle_caf  <- compute_angle_forces.f90:50-93     (caf / caf_d)
 *   oracle_cdf  <- compute_dihedral_forces.f90:24-136 (cdf / cdf_d)
 *   cosine_series, reconstruct <- dipole_reconstruction.f90:37-220
 *
 * `single != 0` reproduces the real(4) variants: position differences are
 * formed in float before promotion (compute_bond_forces.f90:49,
 * compute_angle_forces.f90:58-59, compute_dihedral_forces.f90:35-37), every
 * force / dipole / matrix store rounds to float, and `phi` is left stale when
 * cos_phi == sin_phi == 0 (compute_dihedral_forces.f90:51-55).  `single == 0`
 * is the __double variant.  Arrays are row-major: r, f (n,3); dipoles
 * (D,4,3); transfer_matrix (D,6,3,3) with [m][i][j] = Fortran (m,i,j);
 * coeff (D,6,5) float.  Indices are 0-based int32 as passed by f2py.
 *
 * Divergence, documented: where the reference leaves `reconstruct` outputs
 * undefined (cos^2 gamma >= 1, dipole_reconstruction.f90:108) they are zero
 * here.  The sin(gamma) < 0.1 warning is counted instead of printed.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static long g_clamp_count = 0;
long oracle_clamp_count(void) { return g_clamp_count; }
void oracle_reset_clamp_count(void) { g_clamp_count = 0; }

static inline double fnint(double x) { return round(x); } /* Fortran NINT */
static inline double to_prec(double x, int single) { return single ? (double)(float)x : x; }

static void diff3(const double *r, int32_t i, int32_t j, int single, double out[3]) {
    /* r(i,:) - r(j,:) */
    for (int d = 0; d < 3; ++d) {
        if (single)
            out[d] = (double)((float)r[3 * (int64_t)i + d] - (float)r[3 * (int64_t)j + d]);
        else
            out[d] = r[3 * (int64_t)i + d] - r[3 * (int64_t)j + d];
    }
}

static void mic3(double v[3], const double box[3]) {
    for (int d = 0; d < 3; ++d) v[d] = v[d] - box[d] * fnint(v[d] / box[d]);
}

static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double norm3(const double a[3]) { return sqrt(dot3(a, a)); }
static void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static void acc_force(double *f, int32_t p, const double v[3], double sign, int single) {
    for (int d = 0; d < 3; ++d) f[3 * (int64_t)p + d] = to_prec(f[3 * (int64_t)p + d] + sign * v[d], single);
}

/* compute_bond_forces.f90:42-61 */
double oracle_cbf(int64_t n, double *f, const double *r, const double *box, int64_t m,
                  const int32_t *a, const int32_t *b, const double *r0, const double *k,
                  double *pr, int single) {
    double energy = 0.0;
    memset(f, 0, sizeof(double) * 3 * n);
    for (int64_t t = 0; t < m; ++t) {
        double rab[3], fa[3];
        diff3(r, b[t], a[t], single, rab);
        mic3(rab, box);
        double rn = norm3(rab);
        double df = k[t] * (rn - r0[t]);
        for (int d = 0; d < 3; ++d) fa[d] = -df * rab[d] / rn;
        acc_force(f, a[t], fa, -1.0, single);
        acc_force(f, b[t], fa, +1.0, single);
        double dr = rn - r0[t];
        energy = energy + 0.5 * k[t] * (dr * dr);
        for (int d = 0; d < 3; ++d) pr[d] = to_prec(pr[d] + fa[d] * rab[d], single);
    }
    return energy;
}

/* compute_angle_forces.f90:50-93 */
double oracle_caf(int64_t n, double *f, const double *r, const double *box, int64_t m,
                  const int32_t *a, const int32_t *b, const int32_t *c, const double *t0,
                  const double *k, double *pr, int single) {
    double energy = 0.0;
    memset(f, 0, sizeof(double) * 3 * n);
    pr[0] = pr[1] = pr[2] = 0.0;
    for (int64_t t = 0; t < m; ++t) {
        double ra[3], rc[3], ea[3], ec[3], fa[3], fc[3];
        diff3(r, a[t], b[t], single, ra);
        diff3(r, c[t], b[t], single, rc);
        mic3(ra, box);
        mic3(rc, box);
        double na = norm3(ra), nc = norm3(rc);
        for (int d = 0; d < 3; ++d) { ea[d] = ra[d] / na; ec[d] = rc[d] / nc; }
        double cosphi = dot3(ea, ec);
        double cosphi2 = cosphi * cosphi;
        if (cosphi2 < 1.0) {
            double theta = acos(cosphi);
            double sinphi = sin(theta);
            double dd = theta - t0[t];
            double ff = k[t] * dd;
            double xrasin = -ff / (na * sinphi);
            double xrcsin = -ff / (nc * sinphi);
            for (int d = 0; d < 3; ++d) {
                fa[d] = (ec[d] - cosphi * ea[d]) * xrasin;
                fc[d] = (ea[d] - cosphi * ec[d]) * xrcsin;
            }
            acc_force(f, a[t], fa, -1.0, single);
            acc_force(f, c[t], fc, -1.0, single);
            for (int q = 0; q < 3; ++q) /* f(bb,:) = f(bb,:) + fa + fc, left to right */
                f[3 * (int64_t)b[t] + q] = to_prec((f[3 * (int64_t)b[t] + q] + fa[q]) + fc[q], single);
            energy = energy + 0.5 * ff * dd;
            for (int d = 0; d < 3; ++d) pr[d] = to_prec(pr[d] - (fa[d] * ra[d]) - (fc[d] * rc[d]), single);
        }
    }
    return energy;
}

/* dipole_reconstruction.f90:37-48 */
static void cosine_series(const double cn[5], const double dn[5], double phi, double *energy, double *dE) {
    for (int i = 0; i < 5; ++i) {
        *energy = *energy + cn[i] * (1.0 + cos(i * phi - dn[i]));
        *dE = *dE - i * cn[i] * sin(i * phi - dn[i]);
    }
}

static void cross_matrix(const double M[3][3], const double v[3], double out[3][3]) {
    for (int i = 0; i < 3; ++i) cross3(M[i], v, out[i]);
}
static void outer3(const double a[3], const double b[3], double out[3][3]) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[i][j] = a[i] * b[j];
}

/* dipole_reconstruction.f90:50-220.  rb is the (possibly float) middle bead.
 * dipole: 2x3 out, tm: 3x3x3 out (only written when dipole_flag != 0). */
static void reconstruct(const double rab[3], const double rb[3], const double rcb[3], const double box[3],
                        const double ck[5], const double dk_[5], double phi, int dipole_flag,
                        double *energy_cbt, double *df_cbt, double fa[3], double fb[3], double fc[3],
                        double dipole[2][3], double tm[3][3][3], int single) {
    const double delta = 0.3;
    const double cos_psi = (double)cosf(1.392947f), sin_psi = (double)sinf(1.392947f);
    double gamma_0 = 1.85 - 0.227 * cos(phi - 0.785);
    double dg = 0.227 * sin(phi - 0.785);
    double k = 0.0, dk = 0.0;
    cosine_series(ck, dk_, phi, &k, &dk);
    double norm_a = norm3(rab), norm_c = norm3(rcb);
    double w[3], v[3];
    for (int d = 0; d < 3; ++d) { w[d] = rab[d] / norm_a; v[d] = rcb[d] / norm_c; }
    double cos_gamma = dot3(w, v);
    double cos2 = cos_gamma * cos_gamma;
    *energy_cbt = 0.0;
    *df_cbt = 0.0;
    for (int d = 0; d < 3; ++d) fa[d] = fb[d] = fc[d] = 0.0;
    if (!(cos2 < 1.0)) return;
    double gamm = acos(cos_gamma);
    double sin_gamma = sqrt(1 - cos2);
    if (sin_gamma < 0.1) {
        g_clamp_count++;
        sin_gamma = (double)0.1f;
    }
    for (int d = 0; d < 3; ++d) {
        fa[d] = (v[d] - cos_gamma * w[d]) / norm_a;
        fc[d] = (w[d] - cos_gamma * v[d]) / norm_c;
        fa[d] = -fa[d] / sin_gamma;
        fc[d] = -fc[d] / sin_gamma;
        fb[d] = -(fa[d] + fc[d]);
    }
    double df_ang = k * (gamm - gamma_0);
    double var_sq = (gamm - gamma_0) * (gamm - gamma_0);
    *energy_cbt = 0.5 * k * var_sq;
    *df_cbt = 0.5 * dk * var_sq - df_ang * dg;
    if (dipole_flag == 0) {
        for (int d = 0; d < 3; ++d) { fa[d] *= df_ang; fb[d] *= df_ang; fc[d] *= df_ang; }
        return;
    }
    double fac = exp((gamm - 1.73) / 0.025);
    double theta = -1.607 * gamm + 0.094 + 1.883 / (1.0 + fac);
    double d_theta = -1.607 - 1.883 / 0.025 * fac / ((1.0 + fac) * (1.0 + fac));
    double cos_theta = cos(theta), sin_theta = sin(theta);
    double nn[3], mm[3], r0[3], dd[3];
    cross3(w, v, nn);
    for (int d = 0; d < 3; ++d) nn[d] /= sin_gamma;
    cross3(nn, v, mm);
    for (int d = 0; d < 3; ++d) {
        r0[d] = rb[d] + 0.5 * rcb[d];
        dd[d] = 0.5 * delta * (cos_psi * v[d] + sin_psi * (cos_theta * nn[d] + sin_theta * mm[d]));
        dipole[0][d] = to_prec(r0[d] + dd[d], single);
        dipole[1][d] = to_prec(r0[d] - dd[d], single);
    }
    for (int p = 0; p < 2; ++p)
        for (int d = 0; d < 3; ++d)
            dipole[p][d] = to_prec(dipole[p][d] - box[d] * fnint(dipole[p][d] / box[d]), single);
    double Vb[3][3], Wb[3][3], Vc[3][3], Wa[3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            Vb[i][j] = v[i] * v[j];
            Wb[i][j] = w[i] * w[j];
            if (i == j) { Vb[i][j] -= 1.0; Wb[i][j] -= 1.0; }
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Vb[i][j] /= norm_c;
            Wb[i][j] /= norm_a;
            Vc[i][j] = -Vb[i][j];
            Wa[i][j] = -Wb[i][j];
        }
    double o_a[3][3], o_b[3][3], o_c[3][3], cWa[3][3], cWb[3][3], cVb[3][3], cVc[3][3];
    double Na[3][3], Nb[3][3], Nc[3][3];
    outer3(fa, nn, o_a);
    outer3(fb, nn, o_b);
    outer3(fc, nn, o_c);
    cross_matrix(Wa, v, cWa);
    cross_matrix(Wb, v, cWb);
    cross_matrix(Vb, w, cVb);
    cross_matrix(Vc, w, cVc);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            Na[i][j] = (cos_gamma * o_a[i][j] + cWa[i][j]) / sin_gamma;
            Nb[i][j] = (cos_gamma * o_b[i][j] + cWb[i][j] - cVb[i][j]) / sin_gamma;
            Nc[i][j] = (cos_gamma * o_c[i][j] - cVc[i][j]) / sin_gamma;
        }
    double Ma[3][3], Mb[3][3], Mc[3][3], t1[3][3], t2[3][3];
    cross_matrix(Na, v, Ma);
    cross_matrix(Nb, v, t1);
    cross_matrix(Vb, nn, t2);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Mb[i][j] = t1[i][j] - t2[i][j];
    cross_matrix(Nc, v, t1);
    cross_matrix(Vc, nn, t2);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) Mc[i][j] = t1[i][j] - t2[i][j];
    double pa[3][3], pb[3][3], pc[3][3], qa[3][3], qb[3][3], qc[3][3];
    outer3(fa, nn, pa);
    outer3(fb, nn, pb);
    outer3(fc, nn, pc);
    outer3(fa, mm, qa);
    outer3(fb, mm, qb);
    outer3(fc, mm, qc);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double FNa = sin_theta * d_theta * pa[i][j], FNb = sin_theta * d_theta * pb[i][j],
                   FNc = sin_theta * d_theta * pc[i][j];
            double FMa = cos_theta * d_theta * qa[i][j], FMb = cos_theta * d_theta * qb[i][j],
                   FMc = cos_theta * d_theta * qc[i][j];
            tm[0][i][j] = to_prec(0.5 * delta * (sin_psi * (cos_theta * Na[i][j] + sin_theta * Ma[i][j] + FNa - FMa)), single);
            tm[1][i][j] = to_prec(0.5 * delta * (cos_psi * Vb[i][j] +
                                                 sin_psi * (cos_theta * Nb[i][j] + sin_theta * Mb[i][j] + FNb - FMb)), single);
            tm[2][i][j] = to_prec(0.5 * delta * (cos_psi * Vc[i][j] +
                                                 sin_psi * (cos_theta * Nc[i][j] + sin_theta * Mc[i][j] + FNc - FMc)), single);
        }
    for (int d = 0; d < 3; ++d) { fa[d] *= df_ang; fb[d] *= df_ang; fc[d] *= df_ang; }
}

/* compute_dihedral_forces.f90:24-136 */
double oracle_cdf(int64_t n, double *force, const double *r, double *dipoles, double *tmat,
                  const double *box, int64_t m, const int32_t *a, const int32_t *b, const int32_t *c,
                  const int32_t *d, const float *coeff, const int32_t *dtype, const int32_t *bb_index,
                  int dipole_flag, int single) {
    double energy = 0.0;
    double phi = 0.0, df_dih = 0.0;
    memset(force, 0, sizeof(double) * 3 * n);
    memset(dipoles, 0, sizeof(double) * 12 * m);
    memset(tmat, 0, sizeof(double) * 54 * m);
    for (int64_t t = 0; t < m; ++t) {
        double f[3], g[3], h[3], v[3], w[3], sc[3], fa[3], fb[3], fc[3], fd[3];
        const float *cf = coeff + 30 * t;
        diff3(r, a[t], b[t], single, f);
        diff3(r, b[t], c[t], single, g);
        diff3(r, d[t], c[t], single, h);
        mic3(f, box);
        mic3(g, box);
        mic3(h, box);
        cross3(f, g, v);
        cross3(h, g, w);
        double v_sq = dot3(v, v), w_sq = dot3(w, w), g_norm = norm3(g);
        double cos_phi = dot3(v, w);
        double sin_phi = dot3(w, f) * g_norm;
        if (!(single && cos_phi == 0.0 && sin_phi == 0.0)) phi = atan2(sin_phi, cos_phi);
        double f_dot_g = dot3(f, g), h_dot_g = dot3(h, g);
        if (dtype[t] == 0 || dtype[t] == 1) {
            double cv[5], dv[5], cc[5], dc[5];
            df_dih = 0.0;
            int nzc = 0, nzd = 0;
            for (int i = 0; i < 5; ++i) {
                cv[i] = cf[0 * 5 + i];
                dv[i] = cf[1 * 5 + i];
                cc[i] = cf[2 * 5 + i];
                dc[i] = cf[3 * 5 + i];
                nzc += (cc[i] == 0.0);
                nzd += (dc[i] == 0.0);
            }
            cosine_series(cv, dv, phi, &energy, &df_dih);
            if (nzc != 5 && nzd != 5) cosine_series(cc, dc, phi, &energy, &df_dih);
        }
        if (dtype[t] == 1) {
            double ck[5], dk[5], e_cbt, df_cbt, rb[3], rc[3], mg[3];
            double dip[2][3], tm[3][3][3];
            for (int i = 0; i < 5; ++i) { ck[i] = cf[4 * 5 + i]; dk[i] = cf[5 * 5 + i]; }
            for (int q = 0; q < 3; ++q) {
                rb[q] = single ? (double)(float)r[3 * (int64_t)b[t] + q] : r[3 * (int64_t)b[t] + q];
                rc[q] = single ? (double)(float)r[3 * (int64_t)c[t] + q] : r[3 * (int64_t)c[t] + q];
                mg[q] = -g[q];
            }
            double *dp = dipoles + 12 * t, *tp = tmat + 54 * t;
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 3; ++q) dip[p][q] = dp[3 * p + q];
            for (int x = 0; x < 3; ++x)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) tm[x][i][j] = tp[9 * x + 3 * i + j];
            reconstruct(f, rb, mg, box, ck, dk, phi, dipole_flag, &e_cbt, &df_cbt, fa, fb, fc, dip, tm, single);
            for (int p = 0; p < 2; ++p)
                for (int q = 0; q < 3; ++q) dp[3 * p + q] = dip[p][q];
            for (int x = 0; x < 3; ++x)
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) tp[9 * x + 3 * i + j] = tm[x][i][j];
            energy = energy + e_cbt;
            df_dih = df_dih + df_cbt;
            acc_force(force, a[t], fa, -1.0, single);
            acc_force(force, b[t], fb, -1.0, single);
            acc_force(force, c[t], fc, -1.0, single);
            if (bb_index[t] == 1) {
                for (int p = 0; p < 2; ++p)
                    for (int q = 0; q < 3; ++q) dip[p][q] = dp[6 + 3 * p + q];
                for (int x = 0; x < 3; ++x)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) tm[x][i][j] = tp[27 + 9 * x + 3 * i + j];
                reconstruct(g, rc, h, box, ck, dk, phi, dipole_flag, &e_cbt, &df_cbt, fb, fc, fd, dip, tm, single);
                for (int p = 0; p < 2; ++p)
                    for (int q = 0; q < 3; ++q) dp[6 + 3 * p + q] = dip[p][q];
                for (int x = 0; x < 3; ++x)
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j) tp[27 + 9 * x + 3 * i + j] = tm[x][i][j];
                energy = energy + e_cbt;
                df_dih = df_dih + df_cbt;
                acc_force(force, b[t], fb, -1.0, single);
                acc_force(force, c[t], fc, -1.0, single);
                acc_force(force, d[t], fd, -1.0, single);
            }
        }
        if (dtype[t] == 2) {
            double eq = cf[0], kf = cf[1];
            df_dih = kf * (phi - eq);
            energy = energy + 0.5 * kf * ((phi - eq) * (phi - eq));
        }
        for (int q = 0; q < 3; ++q) {
            sc[q] = v[q] * f_dot_g / (v_sq * g_norm) - w[q] * h_dot_g / (w_sq * g_norm);
            fa[q] = -df_dih * g_norm * v[q] / v_sq;
            fd[q] = df_dih * g_norm * w[q] / w_sq;
            fb[q] = df_dih * sc[q] - fa[q];
            fc[q] = -df_dih * sc[q] - fd[q];
        }
        acc_force(force, a[t], fa, +1.0, single);
        acc_force(force, b[t], fb, +1.0, single);
        acc_force(force, c[t], fc, +1.0, single);
        acc_force(force, d[t], fd, +1.0, single);
    }
    return energy;
}

/* force.py:855-877: project dipole-point forces through the transfer matrices */
void oracle_dipole_redistribution(int64_t n, double *f_on_bead, const double *f_dipoles, const double *tmat,
                                  int64_t m, const int32_t *a, const int32_t *b, const int32_t *c,
                                  const int32_t *d, const int32_t *dtype, const int32_t *last) {
    memset(f_on_bead, 0, sizeof(double) * 3 * n);
    for (int64_t t = 0; t < m; ++t) {
        if (dtype[t] != 1) continue;
        const double *fd = f_dipoles + 12 * t;
        const double *M = tmat + 54 * t;
        int32_t atoms1[3] = {a[t], b[t], c[t]}, atoms2[3] = {b[t], c[t], d[t]};
        for (int pair = 0; pair < (last[t] == 1 ? 2 : 1); ++pair) {
            const double *f0 = fd + 6 * pair, *f1 = fd + 6 * pair + 3;
            const int32_t *atoms = pair ? atoms2 : atoms1;
            double sum[3], diff[3];
            for (int q = 0; q < 3; ++q) { sum[q] = f0[q] + f1[q]; diff[q] = f0[q] - f1[q]; }
            for (int x = 0; x < 3; ++x) {
                const double *Mx = M + 27 * pair + 9 * x;
                for (int i = 0; i < 3; ++i) {
                    double s = Mx[3 * i] * diff[0] + Mx[3 * i + 1] * diff[1] + Mx[3 * i + 2] * diff[2];
                    if (x > 0) s += 0.5 * sum[i];
                    f_on_bead[3 * (int64_t)atoms[x] + i] += s;
                }
            }
        }
    }
}
