/*
 * qsr_oracle.c -- FP64 CPU restatement of the QSR LdG step, energy and
 * observables.  TEST INFRASTRUCTURE ONLY (the checker): it is never linked into
 * or called by the product library; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it.
 *
 * Pinning (what keeps this restatement honest) is listed per function group in
 * the test files tests/test_oracle_*.py: QSR.cuh:160-296 transcripts, golden
 * vectors produced by the reference's own QSR.cpp functions (oracle/_ref),
 * closed forms, invariants and finite-difference variational checks.
 *
 * Loop structure follows QSR.cpp (separate AoS passes, OpenMP over the outer
 * index), arithmetic follows QSR.cu semantics.
 */
#include "qsr_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

#define IDX(i, j, k) ((size_t)(k) + (size_t)Nz * ((size_t)(j) + (size_t)Ny * (size_t)(i)))
#define QC(ptr, id, c) ((ptr)[5 * (size_t)(id) + (c)])

static inline double dmin3(double a, double b, double c) { double m = a < b ? a : b; return m < c ? m : c; }
static inline double dmax2(double a, double b) { return a > b ? a : b; }
static inline int imin3(int a, int b, int c) { int m = a < b ? a : b; return m < c ? m : c; }

/* ------------------------------------------------------------------------- */
/* Geometry: QSR.cu:70-185                                                    */
/* ------------------------------------------------------------------------- */

/* makeDropletGeometry (QSR.cu:144-162).  Extension: radius_index > 0 overrides
 * the inscribed radius (0.5*(min N - 1)), finite c*_index override the centre
 * ((N-1)/2).  With the defaults the reference geometry is reproduced exactly. */
void orc_make_geometry(int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double radius_index, double cx_index, double cy_index, double cz_index,
                       orc_geom* g) {
    g->dx = dx; g->dy = dy; g->dz = dz;
    g->center_x = (isfinite(cx_index) ? cx_index : 0.5 * ((double)Nx - 1.0)) * dx;
    g->center_y = (isfinite(cy_index) ? cy_index : 0.5 * ((double)Ny - 1.0)) * dy;
    g->center_z = (isfinite(cz_index) ? cz_index : 0.5 * ((double)Nz - 1.0)) * dz;
    double ri = (radius_index > 0.0) ? radius_index : 0.5 * ((double)imin3(Nx, Ny, Nz) - 1.0);
    g->radius_x = dmax2(ri * dx, 0.5 * dx);
    g->radius_y = dmax2(ri * dy, 0.5 * dy);
    g->radius_z = dmax2(ri * dz, 0.5 * dz);
    g->min_spacing = dmin3(dx, dy, dz);
    g->min_radius = dmin3(g->radius_x, g->radius_y, g->radius_z);
    g->shell_half_thickness = g->min_spacing;
    g->core_radius = g->min_spacing;
}

/* sampleDropletGeometry (QSR.cu:70-98): out = {x, y, z, center_radius, signed_distance} */
void orc_sample(const orc_geom* g, int i, int j, int k, double out[5]) {
    const double x = i * g->dx - g->center_x;
    const double y = j * g->dy - g->center_y;
    const double z = k * g->dz - g->center_z;
    const double r = sqrt(x * x + y * y + z * z);
    out[0] = x; out[1] = y; out[2] = z; out[3] = r;
    const double sx = x / g->radius_x, sy = y / g->radius_y, sz = z / g->radius_z;
    const double rho_sq = sx * sx + sy * sy + sz * sz;
    if (rho_sq <= 1e-30) { out[4] = -g->min_radius; return; }
    const double rho = sqrt(rho_sq);
    const double gx = x / (g->radius_x * g->radius_x);
    const double gy = y / (g->radius_y * g->radius_y);
    const double gz = z / (g->radius_z * g->radius_z);
    const double gn = sqrt(gx * gx + gy * gy + gz * gz) / rho;
    out[4] = (gn > 1e-30) ? (rho - 1.0) / gn : r - g->min_radius;
}

/* getCenterDirection / getSurfaceNormal (QSR.cu:117-142) */
static int center_direction(const double s[5], double n[3]) {
    if (s[3] <= 1e-30) { n[0] = 1.0; n[1] = 0.0; n[2] = 0.0; return 0; }
    n[0] = s[0] / s[3]; n[1] = s[1] / s[3]; n[2] = s[2] / s[3];
    return 1;
}

static int surface_normal_s(const orc_geom* g, const double s[5], double n[3]) {
    const double gx = s[0] / (g->radius_x * g->radius_x);
    const double gy = s[1] / (g->radius_y * g->radius_y);
    const double gz = s[2] / (g->radius_z * g->radius_z);
    const double gn = sqrt(gx * gx + gy * gy + gz * gz);
    if (gn <= 1e-30) return center_direction(s, n);
    n[0] = gx / gn; n[1] = gy / gn; n[2] = gz / gn;
    return 1;
}

int orc_surface_normal(const orc_geom* g, int i, int j, int k, double n[3]) {
    double s[5];
    orc_sample(g, i, j, k, s);
    return surface_normal_s(g, s, n);
}

/* make_uniaxial_qtensor (QSR.cu:60-68) */
static void uniaxial(double S, const double n[3], double q[5]) {
    q[0] = S * (n[0] * n[0] - 1.0 / 3.0);
    q[1] = S * (n[0] * n[1]);
    q[2] = S * (n[0] * n[2]);
    q[3] = S * (n[1] * n[1] - 1.0 / 3.0);
    q[4] = S * (n[1] * n[2]);
}

/* shell mask build (QSR.cu:2160-2172) with dropletIsInShell (QSR.cu:104-106) */
size_t orc_shell_mask(const orc_geom* g, int Nx, int Ny, int Nz, uint8_t* shell) {
    size_t count = 0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                double s[5];
                orc_sample(g, i, j, k, s);
                const int in = fabs(s[4]) <= g->shell_half_thickness;
                shell[IDX(i, j, k)] = (uint8_t)in;
                count += (size_t)in;
            }
    return count;
}

/* estimateEllipsoidSurfaceArea (QSR.cu:164-172): Thomsen p = 1.6075 */
double orc_surface_area(const orc_geom* g) {
    const double p = 1.6075;
    const double ap = pow(g->radius_x, p), bp = pow(g->radius_y, p), cp = pow(g->radius_z, p);
    const double mean_term = (ap * bp + ap * cp + bp * cp) / 3.0;
    if (!(mean_term > 0.0)) return NAN;
    return 4.0 * ORC_PI * pow(mean_term, 1.0 / p);
}

/* estimateShellThicknessFromMask (QSR.cu:174-185) */
double orc_shell_thickness(const orc_geom* g, size_t shell_count) {
    const double nominal = 2.0 * g->shell_half_thickness;
    if (shell_count == 0) return nominal;
    const double vol = (double)shell_count * g->dx * g->dy * g->dz;
    const double area = orc_surface_area(g);
    if (!(area > 0.0)) return nominal;
    const double t = vol / area;
    if (!isfinite(t) || !(t > 0.0)) return nominal;
    return t;
}

/* ------------------------------------------------------------------------- */
/* Bulk constants, S_eq, transitions, shell order, dt (QSR.cu:1363-1604)      */
/* ------------------------------------------------------------------------- */

/* bulk_ABC_from_convention (QSR.cu:1363-1378) */
void orc_bulk_ABC(double a, double b, double c, double T, double T_star, int mode, double out[3]) {
    const double dT = T - T_star;
    out[0] = 3.0 * a * dT;
    if (mode == 1) { out[1] = (9.0 * b) / 2.0; out[2] = (9.0 * c) / 2.0; }
    else { out[1] = (27.0 * b) / 2.0; out[2] = 9.0 * c; }
}

/* bulk_fS_uniaxial (QSR.cu:1380-1389) */
double orc_bulk_fS(double A, double B, double C, double S) {
    const double S2 = S * S;
    return (A / 3.0) * S2 - (2.0 * B / 27.0) * (S2 * S) + (C / 9.0) * (S2 * S2);
}

/* S_eq_uniaxial_from_ABC (QSR.cu:1391-1418) */
double orc_S_eq(double A, double B, double C) {
    if (!(C > 0.0)) return 0.0;
    const double disc = B * B - 24.0 * A * C;
    if (!(disc > 0.0)) return 0.0;
    const double sd = sqrt(disc);
    const double roots[2] = {(B - sd) / (4.0 * C), (B + sd) / (4.0 * C)};
    double best_S = 0.0, best_f = 0.0;
    for (int r = 0; r < 2; ++r) {
        const double S = roots[r];
        if (!(S > 0.0)) continue;
        const double f = orc_bulk_fS(A, B, C, S);
        if (f < best_f) { best_f = f; best_S = S; }
    }
    return best_S;
}

/* QSR.cu:1420-1438 */
double orc_A_NI(double B, double C) { return (C > 0.0) ? (B * B) / (27.0 * C) : NAN; }
double orc_A_spinN(double B, double C) { return (C > 0.0) ? (B * B) / (24.0 * C) : NAN; }
double orc_T_from_A(double A, double a, double T_star) { return (a > 0.0) ? T_star + A / (3.0 * a) : NAN; }

/* equilibrium_shell_order_from_temperature / resolve_boundary_shell_order (QSR.cu:1461-1487) */
double orc_shell_order(int custom, double custom_S, double a, double b, double c,
                       double T, double T_star, int mode) {
    if (custom) return custom_S;
    double abc[3];
    orc_bulk_ABC(a, b, c, T, T_star, mode, abc);
    const double s = orc_S_eq(abc[0], abc[1], abc[2]);
    return (s > 0.0) ? s : 0.0;
}

/* estimate_stable_dt_window (QSR.cu:1528-1604) */
void orc_dt_estimate(double dx, double dy, double dz, double gamma, double a, double b, double c,
                     double T_star, int mode, double kappa, double L1, double L2, double L3,
                     double S_ref_used, double S0, double T_a, double T_b,
                     int use_semi_implicit, double L_stab, int jacobi_iters, double out[10]) {
    const double min_dx = dmin3(dx, dy, dz);
    double S_scale = isfinite(S_ref_used) ? fabs(S_ref_used) : dmax2(fabs(S0), 0.5);
    const double Sa = fabs(orc_shell_order(0, 0.0, a, b, c, T_a, T_star, mode));
    const double Sb = fabs(orc_shell_order(0, 0.0, a, b, c, T_b, T_star, mode));
    if (isfinite(Sa) && Sa > 0.0) S_scale = dmax2(S_scale, Sa);
    if (isfinite(Sb) && Sb > 0.0) S_scale = dmax2(S_scale, Sb);
    S_scale = dmax2(S_scale, 1.0);
    const int ldg = (L1 != 0.0 || L2 != 0.0 || L3 != 0.0);
    const int aniso = (L2 != 0.0 || L3 != 0.0);
    double L_eff_total;
    if (ldg) {
        L_eff_total = fabs(L1);
        L_eff_total = dmax2(L_eff_total, fabs(L2));
        L_eff_total = dmax2(L_eff_total, fabs(L1 + L2));
        L_eff_total = dmax2(L_eff_total, fabs(L1 + 2.0 * L2));
        L_eff_total = dmax2(L_eff_total, fabs(L3) * S_scale);
    } else {
        L_eff_total = fabs(kappa);
    }
    if (!(L_eff_total > 0.0)) L_eff_total = 1e-12;
    double D_total = L_eff_total / gamma;
    if (!(D_total > 0.0) || !isfinite(D_total)) D_total = 1e-12;
    double dt_diff_total = 0.5 * (min_dx * min_dx) / (6.0 * D_total);
    if (aniso) dt_diff_total *= 0.25;
    double L_eff_explicit = L_eff_total;
    double dt_diff = dt_diff_total;
    if (use_semi_implicit && L_stab > 0.0 && jacobi_iters > 0) {
        const double L_base = ldg ? L1 : kappa;
        const double L_stab_signed = (L_base >= 0.0) ? L_stab : -L_stab;
        const double L_rem = fabs(L_base - L_stab_signed);
        if (ldg) {
            L_eff_explicit = L_rem;
            L_eff_explicit = dmax2(L_eff_explicit, fabs(L2));
            L_eff_explicit = dmax2(L_eff_explicit, fabs(L1 + L2));
            L_eff_explicit = dmax2(L_eff_explicit, fabs(L1 + 2.0 * L2));
            L_eff_explicit = dmax2(L_eff_explicit, fabs(L3) * S_scale);
        } else {
            L_eff_explicit = L_rem;
        }
        if (L_eff_explicit > 0.0) {
            dt_diff = 0.5 * (min_dx * min_dx) / (6.0 * (L_eff_explicit / gamma));
            if (aniso) dt_diff *= 0.25;
        } else {
            dt_diff = INFINITY;
        }
    }
    double rate = 0.0;
    const double Ts[2] = {T_a, T_b};
    for (int t = 0; t < 2; ++t) {
        double abc[3];
        orc_bulk_ABC(a, b, c, Ts[t], T_star, mode, abc);
        const double r = fabs(abc[0]) + fabs(abc[1]) * S_scale + fabs(abc[2]) * (S_scale * S_scale);
        rate = (t == 0) ? r : dmax2(rate, r);
    }
    const double dt_bulk = (rate > 0.0) ? (0.02 * gamma / rate) : INFINITY;
    double dt_max = dt_diff < dt_bulk ? dt_diff : dt_bulk;
    if (!(dt_max > 0.0) || !isfinite(dt_max)) dt_max = dt_diff;
    out[0] = min_dx; out[1] = S_scale; out[2] = L_eff_total; out[3] = D_total;
    out[4] = dt_diff_total; out[5] = L_eff_explicit; out[6] = dt_diff; out[7] = rate;
    out[8] = dt_bulk; out[9] = dt_max;
}

/* ------------------------------------------------------------------------- */
/* Step kernels: QSR.cu:279-867 (loop structure of QSR.cpp:179-918)           */
/* ------------------------------------------------------------------------- */

/* computeLaplacianKernel (QSR.cu:279-323): one-sided second difference on faces */
void orc_laplacian(const double* Q, double* lap, int Nx, int Ny, int Nz, double dx, double dy, double dz) {
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                size_t xa, xb, xc, ya, yb, yc, za, zb, zc; /* a - 2b + c */
                if (i > 0 && i < Nx - 1) { xa = IDX(i + 1, j, k); xb = id; xc = IDX(i - 1, j, k); }
                else if (i == 0) { xa = IDX(i + 2, j, k); xb = IDX(i + 1, j, k); xc = id; }
                else { xa = id; xb = IDX(i - 1, j, k); xc = IDX(i - 2, j, k); }
                if (j > 0 && j < Ny - 1) { ya = IDX(i, j + 1, k); yb = id; yc = IDX(i, j - 1, k); }
                else if (j == 0) { ya = IDX(i, j + 2, k); yb = IDX(i, j + 1, k); yc = id; }
                else { ya = id; yb = IDX(i, j - 1, k); yc = IDX(i, j - 2, k); }
                if (k > 0 && k < Nz - 1) { za = IDX(i, j, k + 1); zb = id; zc = IDX(i, j, k - 1); }
                else if (k == 0) { za = IDX(i, j, k + 2); zb = IDX(i, j, k + 1); zc = id; }
                else { za = id; zb = IDX(i, j, k - 1); zc = IDX(i, j, k - 2); }
                for (int c = 0; c < 5; ++c) {
                    const double d2x = (QC(Q, xa, c) - QC(Q, xb, c) * 2.0 + QC(Q, xc, c)) * ix2;
                    const double d2y = (QC(Q, ya, c) - QC(Q, yb, c) * 2.0 + QC(Q, yc, c)) * iy2;
                    const double d2z = (QC(Q, za, c) - QC(Q, zb, c) * 2.0 + QC(Q, zc, c)) * iz2;
                    QC(lap, id, c) = d2x + d2y + d2z;
                }
            }
}

/* get_Q_deriv_device (QSR.cu:255-273); comp 0..4 = Qxx,Qxy,Qxz,Qyy,Qyz */
static inline double qderiv(const double* Q, size_t id, size_t stride, int n, int cur, double inv_2d,
                            double inv_d, int comp) {
    if (cur > 0 && cur < n - 1) return (QC(Q, id + stride, comp) - QC(Q, id - stride, comp)) * inv_2d;
    if (cur == 0) return (QC(Q, id + stride, comp) - QC(Q, id, comp)) * inv_d;
    return (QC(Q, id, comp) - QC(Q, id - stride, comp)) * inv_d;
}

/* get_deriv_device (QSR.cu:245-253) */
static inline double sderiv(const double* D, size_t id, size_t stride, int n, int cur, double inv_2d,
                            double inv_d) {
    if (cur > 0 && cur < n - 1) return (D[id + stride] - D[id - stride]) * inv_2d;
    if (cur == 0) return (D[id + stride] - D[id]) * inv_d;
    return (D[id] - D[id - stride]) * inv_d;
}

/* computeDivergenceKernel (QSR.cu:325-371): Dcol_j = d_k Q_kj */
void orc_divergence(const double* Q, double* Dx, double* Dy, double* Dz,
                    int Nx, int Ny, int Nz, double dx, double dy, double dz) {
    const double i2x = 1.0 / (2.0 * dx), i2y = 1.0 / (2.0 * dy), i2z = 1.0 / (2.0 * dz);
    const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
    const size_t sx = (size_t)Ny * Nz, sy = (size_t)Nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                Dx[id] = qderiv(Q, id, sx, Nx, i, i2x, ix, 0) + qderiv(Q, id, sy, Ny, j, i2y, iy, 1) +
                         qderiv(Q, id, 1, Nz, k, i2z, iz, 2);
                Dy[id] = qderiv(Q, id, sx, Nx, i, i2x, ix, 1) + qderiv(Q, id, sy, Ny, j, i2y, iy, 3) +
                         qderiv(Q, id, 1, Nz, k, i2z, iz, 4);
                const double a = qderiv(Q, id, sx, Nx, i, i2x, ix, 2);
                const double b = qderiv(Q, id, sy, Ny, j, i2y, iy, 4);
                const double zc = -(QC(Q, id, 0) + QC(Q, id, 3));
                double dzz;
                if (k > 0 && k < Nz - 1) {
                    const double zp = -(QC(Q, id + 1, 0) + QC(Q, id + 1, 3));
                    const double zm = -(QC(Q, id - 1, 0) + QC(Q, id - 1, 3));
                    dzz = (zp - zm) * i2z;
                } else if (k == 0) {
                    const double zp = -(QC(Q, id + 1, 0) + QC(Q, id + 1, 3));
                    dzz = (zp - zc) * iz;
                } else {
                    const double zm = -(QC(Q, id - 1, 0) + QC(Q, id - 1, 3));
                    dzz = (zc - zm) * iz;
                }
                Dz[id] = a + b + dzz;
            }
}

/* bulk molecular-field coefficients (QSR.cu:385-397) */
static void bulk_force_coeffs(const orc_phys* p, double* cQ, double* cQ2, double* cQ3) {
    const double dT = p->T - p->T_star;
    *cQ = 3.0 * p->a * dT;
    if (p->modechoice == 1) { *cQ2 = (9.0 * p->b) / 2.0; *cQ3 = (9.0 * p->c) / 2.0; }
    else { *cQ2 = (27.0 * p->b) / 2.0; *cQ3 = 9.0 * p->c; }
}

/* computeChemicalPotentialKernel (QSR.cu:373-472) */
void orc_chempot(const double* Q, double* mu, const double* lap, int Nx, int Ny, int Nz,
                 double dx, double dy, double dz, const orc_phys* p,
                 const double* Dx, const double* Dy, const double* Dz) {
    double cQ, cQ2, cQ3;
    bulk_force_coeffs(p, &cQ, &cQ2, &cQ3);
    const double i2x = 1.0 / (2.0 * dx), i2y = 1.0 / (2.0 * dy), i2z = 1.0 / (2.0 * dz);
    const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
    const size_t sx = (size_t)Ny * Nz, sy = (size_t)Nz;
    const int no_L = (p->L1 == 0.0 && p->L2 == 0.0 && p->L3 == 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                const double xx = QC(Q, id, 0), xy = QC(Q, id, 1), xz = QC(Q, id, 2);
                const double yy = QC(Q, id, 3), yz = QC(Q, id, 4), zz = -xx - yy;
                const double tr2 = xx * xx + yy * yy + zz * zz + 2 * (xy * xy + xz * xz + yz * yz);
                const double q2xx = xx * xx + xy * xy + xz * xz;
                const double q2xy = xx * xy + xy * yy + xz * yz;
                const double q2xz = xx * xz + xy * yz + xz * zz;
                const double q2yy = xy * xy + yy * yy + yz * yz;
                const double q2yz = xy * xz + yy * yz + yz * zz;
                const double q2zz = xz * xz + yz * yz + zz * zz;
                double hxx = cQ * xx - cQ2 * (q2xx - tr2 / 3.0) + cQ3 * tr2 * xx;
                double hyy = cQ * yy - cQ2 * (q2yy - tr2 / 3.0) + cQ3 * tr2 * yy;
                double hzz = cQ * zz - cQ2 * (q2zz - tr2 / 3.0) + cQ3 * tr2 * zz;
                double hxy = cQ * xy - cQ2 * q2xy + cQ3 * tr2 * xy;
                double hxz = cQ * xz - cQ2 * q2xz + cQ3 * tr2 * xz;
                double hyz = cQ * yz - cQ2 * q2yz + cQ3 * tr2 * yz;
                const double* L = lap + 5 * id;
                if (no_L) {
                    hxx -= p->kappa * L[0]; hyy -= p->kappa * L[3]; hzz -= p->kappa * (-L[0] - L[3]);
                    hxy -= p->kappa * L[1]; hxz -= p->kappa * L[2]; hyz -= p->kappa * L[4];
                }
                if (p->L1 != 0.0) {
                    hxx -= p->L1 * L[0]; hyy -= p->L1 * L[3]; hzz -= p->L1 * (-L[0] - L[3]);
                    hxy -= p->L1 * L[1]; hxz -= p->L1 * L[2]; hyz -= p->L1 * L[4];
                }
                if (p->L2 != 0.0) {
                    const double dDx_dx = sderiv(Dx, id, sx, Nx, i, i2x, ix);
                    const double dDy_dy = sderiv(Dy, id, sy, Ny, j, i2y, iy);
                    const double dDz_dz = sderiv(Dz, id, 1, Nz, k, i2z, iz);
                    const double dDx_dy = sderiv(Dx, id, sy, Ny, j, i2y, iy);
                    const double dDy_dx = sderiv(Dy, id, sx, Nx, i, i2x, ix);
                    const double dDx_dz = sderiv(Dx, id, 1, Nz, k, i2z, iz);
                    const double dDz_dx = sderiv(Dz, id, sx, Nx, i, i2x, ix);
                    const double dDy_dz = sderiv(Dy, id, 1, Nz, k, i2z, iz);
                    const double dDz_dy = sderiv(Dz, id, sy, Ny, j, i2y, iy);
                    hxx -= p->L2 * dDx_dx; hyy -= p->L2 * dDy_dy; hzz -= p->L2 * dDz_dz;
                    hxy -= p->L2 * (dDy_dx + dDx_dy);
                    hxz -= p->L2 * (dDz_dx + dDx_dz);
                    hyz -= p->L2 * (dDz_dy + dDy_dz);
                }
                const double corr = (hxx + hyy + hzz) / 3.0;
                double* h = mu + 5 * id;
                h[0] = hxx - corr; h[3] = hyy - corr; h[1] = hxy; h[2] = hxz; h[4] = hyz;
            }
}

/* computeChemicalPotentialL3Kernel (QSR.cu:474-621): interior only, adds to the
 * already-projected mu and re-projects. */
void orc_chempot_l3(const double* Q, double* mu, int Nx, int Ny, int Nz,
                    double dx, double dy, double dz, const orc_phys* p,
                    const double* Dx, const double* Dy, const double* Dz) {
    if (p->L3 == 0.0) return;
    const double L3 = p->L3;
    const double i2x = 1.0 / (2.0 * dx), i2y = 1.0 / (2.0 * dy), i2z = 1.0 / (2.0 * dz);
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
    const double i4xy = 1.0 / (4.0 * dx * dy), i4xz = 1.0 / (4.0 * dx * dz), i4yz = 1.0 / (4.0 * dy * dz);
    const size_t sx = (size_t)Ny * Nz, sy = (size_t)Nz, sz = 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < Nx - 1; ++i)
        for (int j = 1; j < Ny - 1; ++j)
            for (int k = 1; k < Nz - 1; ++k) {
                const size_t id = IDX(i, j, k);
                const double qxx = QC(Q, id, 0), qxy = QC(Q, id, 1), qxz = QC(Q, id, 2);
                const double qyy = QC(Q, id, 3), qyz = QC(Q, id, 4), qzz = -qxx - qyy;
                double* h = mu + 5 * id;
                double hxx = h[0], hyy = h[3], hzz = -(h[0] + h[3]), hxy = h[1], hxz = h[2], hyz = h[4];
                /* first derivatives d[c][dir], comps 0..4, central (interior) */
                double d[6][3];
                for (int c = 0; c < 5; ++c) {
                    d[c][0] = (QC(Q, id + sx, c) - QC(Q, id - sx, c)) * i2x;
                    d[c][1] = (QC(Q, id + sy, c) - QC(Q, id - sy, c)) * i2y;
                    d[c][2] = (QC(Q, id + sz, c) - QC(Q, id - sz, c)) * i2z;
                }
                for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
                /* M_ab = sum_kl dQkl/da dQkl/db with off-diagonals counted twice */
                double M[3][3];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b)
                        M[a][b] = d[0][a] * d[0][b] + d[3][a] * d[3][b] + d[5][a] * d[5][b] +
                                  2.0 * (d[1][a] * d[1][b] + d[2][a] * d[2][b] + d[4][a] * d[4][b]);
                hxx += 0.5 * L3 * M[0][0]; hyy += 0.5 * L3 * M[1][1]; hzz += 0.5 * L3 * M[2][2];
                hxy += 0.5 * L3 * M[0][1]; hxz += 0.5 * L3 * M[0][2]; hyz += 0.5 * L3 * M[1][2];
                /* B(comp) = Dcol . grad Qcomp + Q_ij d_i d_j Qcomp (QSR.cu:576-594) */
                double Bv[5];
                static const int comps[5] = {0, 3, 1, 2, 4};
                for (int t = 0; t < 5; ++t) {
                    const int c = comps[t];
                    const double Qc = QC(Q, id, c);
                    const double d2xx = (QC(Q, id + sx, c) - 2.0 * Qc + QC(Q, id - sx, c)) * ix2;
                    const double d2yy = (QC(Q, id + sy, c) - 2.0 * Qc + QC(Q, id - sy, c)) * iy2;
                    const double d2zz = (QC(Q, id + sz, c) - 2.0 * Qc + QC(Q, id - sz, c)) * iz2;
                    const double d2xy = (QC(Q, id + sx + sy, c) - QC(Q, id + sx - sy, c) -
                                         QC(Q, id - sx + sy, c) + QC(Q, id - sx - sy, c)) * i4xy;
                    const double d2xz = (QC(Q, id + sx + sz, c) - QC(Q, id + sx - sz, c) -
                                         QC(Q, id - sx + sz, c) + QC(Q, id - sx - sz, c)) * i4xz;
                    const double d2yz = (QC(Q, id + sy + sz, c) - QC(Q, id + sy - sz, c) -
                                         QC(Q, id - sy + sz, c) + QC(Q, id - sy - sz, c)) * i4yz;
                    const double contract = qxx * d2xx + qyy * d2yy + qzz * d2zz +
                                            2.0 * (qxy * d2xy + qxz * d2xz + qyz * d2yz);
                    Bv[c] = Dx[id] * d[c][0] + Dy[id] * d[c][1] + Dz[id] * d[c][2] + contract;
                }
                const double Bzz = -(Bv[0] + Bv[3]);
                hxx -= L3 * Bv[0]; hyy -= L3 * Bv[3]; hzz -= L3 * Bzz;
                hxy -= L3 * Bv[1]; hxz -= L3 * Bv[2]; hyz -= L3 * Bv[4];
                const double corr = (hxx + hyy + hzz) / 3.0;
                h[0] = hxx - corr; h[3] = hyy - corr; h[1] = hxy; h[2] = hxz; h[4] = hyz;
            }
}

/* applyWeakAnchoringPenaltyKernel (QSR.cu:623-652): mu += (W/delta)(Q - Q0) on shell */
void orc_weak_anchor(double* mu, const double* Q, int Nx, int Ny, int Nz, const uint8_t* shell,
                     const orc_geom* g, double S_shell, double W, double thickness) {
    if (W == 0.0 || !(thickness > 0.0)) return;
    const double W_eff = W / thickness;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                if (!shell[id]) continue;
                double n[3], q0[5];
                orc_surface_normal(g, i, j, k, n);
                uniaxial(S_shell, n, q0);
                for (int c = 0; c < 5; ++c) QC(mu, id, c) += W_eff * (QC(Q, id, c) - q0[c]);
            }
}

/* qtensor_frobenius_norm (QSR.cu:26-32) */
static double frob(const double* q) {
    const double xx = q[0], xy = q[1], xz = q[2], yy = q[3], yz = q[4], zz = -xx - yy;
    const double t = xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz);
    return sqrt(t > 0.0 ? t : 0.0);
}

/* updateQTensorKernel (QSR.cu:697-732) with dropletShouldEvolve (QSR.cu:108-111) */
void orc_update(double* Q, const double* mu, int Nx, int Ny, int Nz, double dt, const uint8_t* shell,
                const orc_geom* g, double gamma, double W, double q_norm_cap) {
    const double mob = 1.0 / gamma;
    const double band = (W > 0.0) ? g->shell_half_thickness : 0.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                if (W == 0.0 && shell[id]) continue;
                double s[5];
                orc_sample(g, i, j, k, s);
                if (!(s[4] <= band)) continue;
                double* q = Q + 5 * id;
                const double* h = mu + 5 * id;
                for (int c = 0; c < 5; ++c) q[c] -= mob * h[c] * dt;
                if (q_norm_cap > 0.0) {
                    const double nrm = frob(q);
                    if (nrm > q_norm_cap && nrm > 0.0) {
                        const double sc = q_norm_cap / nrm;
                        for (int c = 0; c < 5; ++c) q[c] *= sc;
                    }
                }
            }
}

/* applyBoundaryConditionsKernel (QSR.cu:851-867) */
void orc_bc(double* Q, int Nx, int Ny, int Nz, const uint8_t* shell, const orc_geom* g,
            double S_shell, double W) {
    if (W > 0.0) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                if (!shell[id]) continue;
                double n[3];
                orc_surface_normal(g, i, j, k, n);
                uniaxial(S_shell, n, Q + 5 * id);
            }
}

/* computeRhsSemiImplicitKernel (QSR.cu:734-784) */
void orc_rhs_semi_implicit(const double* Q, const double* mu, const double* lap, double* rhs,
                           int Nx, int Ny, int Nz, double dt, double gamma, double L_stab,
                           const uint8_t* shell, const orc_geom* g, double W) {
    const double mob = 1.0 / gamma, alpha = dt * mob * L_stab;
    const double band = (W > 0.0) ? g->shell_half_thickness : 0.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                double s[5];
                orc_sample(g, i, j, k, s);
                if ((W == 0.0 && shell[id]) || !(s[4] <= band)) {
                    for (int c = 0; c < 5; ++c) QC(rhs, id, c) = QC(Q, id, c);
                    continue;
                }
                for (int c = 0; c < 5; ++c)
                    QC(rhs, id, c) = QC(Q, id, c) - (dt * mob) * QC(mu, id, c) - alpha * QC(lap, id, c);
            }
}

/* jacobiHelmholtzStepKernel (QSR.cu:786-849) */
void orc_jacobi_step(const double* rhs, const double* Qold, double* Qnew, int Nx, int Ny, int Nz,
                     double alpha, double dx, double dy, double dz, const uint8_t* shell,
                     const orc_geom* g, double W) {
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
    const double denom = 1.0 + 2.0 * alpha * (ix2 + iy2 + iz2);
    const double band = (W > 0.0) ? g->shell_half_thickness : 0.0;
    const size_t sx = (size_t)Ny * Nz, sy = (size_t)Nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                int frozen = (i == 0 || j == 0 || k == 0 || i == Nx - 1 || j == Ny - 1 || k == Nz - 1);
                if (!frozen && W == 0.0 && shell[id]) frozen = 1;
                if (!frozen) {
                    double s[5];
                    orc_sample(g, i, j, k, s);
                    if (!(s[4] <= band)) frozen = 1;
                }
                if (!frozen && !(denom > 0.0)) frozen = 1;
                if (frozen) { for (int c = 0; c < 5; ++c) QC(Qnew, id, c) = QC(Qold, id, c); continue; }
                for (int c = 0; c < 5; ++c)
                    QC(Qnew, id, c) = (QC(rhs, id, c) +
                                       alpha * (ix2 * (QC(Qold, id + sx, c) + QC(Qold, id - sx, c)) +
                                                iy2 * (QC(Qold, id + sy, c) + QC(Qold, id - sy, c)) +
                                                iz2 * (QC(Qold, id + 1, c) + QC(Qold, id - 1, c)))) / denom;
            }
}

/* One explicit iteration in the launch order of QSR.cu:3223-3238 */
void orc_step(double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
              const uint8_t* shell, double thickness, double S_shell, double dt, double gamma,
              double q_norm_cap, double* work) {
    const size_t N = (size_t)Nx * Ny * Nz;
    double* lap = work;
    double* mu = work + 5 * N;
    double* Dx = work + 10 * N;
    double* Dy = Dx + N;
    double* Dz = Dy + N;
    orc_laplacian(Q, lap, Nx, Ny, Nz, g->dx, g->dy, g->dz);
    if (p->L2 != 0.0 || p->L3 != 0.0) orc_divergence(Q, Dx, Dy, Dz, Nx, Ny, Nz, g->dx, g->dy, g->dz);
    orc_chempot(Q, mu, lap, Nx, Ny, Nz, g->dx, g->dy, g->dz, p, Dx, Dy, Dz);
    if (p->L3 != 0.0) orc_chempot_l3(Q, mu, Nx, Ny, Nz, g->dx, g->dy, g->dz, p, Dx, Dy, Dz);
    orc_weak_anchor(mu, Q, Nx, Ny, Nz, shell, g, S_shell, p->W, thickness);
    orc_update(Q, mu, Nx, Ny, Nz, dt, shell, g, gamma, p->W, q_norm_cap);
    orc_bc(Q, Nx, Ny, Nz, shell, g, S_shell, p->W);
}

/* One IMEX iteration (QSR.cu:2713-2753, semi-implicit branch) */
void orc_step_imex(double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
                   const uint8_t* shell, double thickness, double S_shell, double dt, double gamma,
                   double L_stab, int jacobi_iters, double* work) {
    const size_t N = (size_t)Nx * Ny * Nz;
    double* lap = work;
    double* mu = work + 5 * N;
    double* Dx = work + 10 * N;
    double* Dy = Dx + N;
    double* Dz = Dy + N;
    double* rhs = work + 13 * N;
    double* alt = work + 18 * N;
    orc_laplacian(Q, lap, Nx, Ny, Nz, g->dx, g->dy, g->dz);
    if (p->L2 != 0.0 || p->L3 != 0.0) orc_divergence(Q, Dx, Dy, Dz, Nx, Ny, Nz, g->dx, g->dy, g->dz);
    orc_chempot(Q, mu, lap, Nx, Ny, Nz, g->dx, g->dy, g->dz, p, Dx, Dy, Dz);
    if (p->L3 != 0.0) orc_chempot_l3(Q, mu, Nx, Ny, Nz, g->dx, g->dy, g->dz, p, Dx, Dy, Dz);
    orc_weak_anchor(mu, Q, Nx, Ny, Nz, shell, g, S_shell, p->W, thickness);
    orc_rhs_semi_implicit(Q, mu, lap, rhs, Nx, Ny, Nz, dt, gamma, L_stab, shell, g, p->W);
    const double alpha = (dt / gamma) * L_stab;
    double* qold = Q;
    double* qnew = alt;
    for (int it = 0; it < jacobi_iters; ++it) {
        orc_jacobi_step(rhs, qold, qnew, Nx, Ny, Nz, alpha, g->dx, g->dy, g->dz, shell, g, p->W);
        double* t = qold; qold = qnew; qnew = t;
    }
    if (qold != Q) memcpy(Q, qold, 5 * N * sizeof(double));
    orc_bc(Q, Nx, Ny, Nz, shell, g, S_shell, p->W);
}

/* ------------------------------------------------------------------------- */
/* Observables                                                                */
/* ------------------------------------------------------------------------- */

/* calculateNematicFieldDevice (QSR.cu:191-243): out = {S, nx, ny, nz} */
void orc_nematic(const double q5[5], double out[4]) {
    const double xx = q5[0], xy = q5[1], xz = q5[2], yy = q5[3], yz = q5[4], zz = -xx - yy;
    if (fabs(xx) < 1e-9 && fabs(xy) < 1e-9 && fabs(xz) < 1e-9 && fabs(yy) < 1e-9 && fabs(yz) < 1e-9) {
        out[0] = 0.0; out[1] = 0.0; out[2] = 0.0; out[3] = 1.0;
        return;
    }
    double nx = xx + xy + xz, ny = xy + yy + yz, nz = xz + yz - xx - yy;
    const double n0 = sqrt(nx * nx + ny * ny + nz * nz);
    if (n0 > 1e-12) { nx /= n0; ny /= n0; nz /= n0; }
    else { nx = 0.0; ny = 0.0; nz = 1.0; }
    const double r1 = fabs(xx) + fabs(xy) + fabs(xz);
    const double r2 = fabs(xy) + fabs(yy) + fabs(yz);
    const double r3 = fabs(xz) + fabs(yz) + fabs(zz);
    const double alpha = fmax(r1, fmax(r2, r3)) + 1e-12;
    for (int it = 0; it < 15; ++it) {
        const double ax = xx * nx + xy * ny + xz * nz + alpha * nx;
        const double ay = xy * nx + yy * ny + yz * nz + alpha * ny;
        const double az = xz * nx + yz * ny + zz * nz + alpha * nz;
        const double nrm = sqrt(ax * ax + ay * ay + az * az);
        if (nrm < 1e-13) break;
        nx = ax / nrm; ny = ay / nrm; nz = az / nrm;
    }
    const double lam = nx * (xx * nx + xy * ny + xz * nz) + ny * (xy * nx + yy * ny + yz * nz) +
                       nz * (xz * nx + yz * ny + zz * nz);
    if (nz < -1e-10) { nx = -nx; ny = -ny; nz = -nz; }
    out[0] = 1.5 * lam; out[1] = nx; out[2] = ny; out[3] = nz;
}

/* computeEnergyKernel (QSR.cu:869-992) + computeAnchoringEnergyKernel (QSR.cu:654-695),
 * summed sequentially in index order like reduce_energy (QSR.cu:2323-2341). */
void orc_energy(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
                const uint8_t* shell, double S_shell, double thickness, double out[3]) {
    const size_t N = (size_t)Nx * Ny * Nz;
    double* eb = (double*)calloc(N, sizeof(double));
    double* ee = (double*)calloc(N, sizeof(double));
    double* ea = (double*)calloc(N, sizeof(double));
    const double dx = g->dx, dy = g->dy, dz = g->dz, dV = dx * dy * dz;
    const double dT = p->T - p->T_star;
    double c2, c3, c4;
    if (p->modechoice == 1) { c2 = (3.0 * p->a * dT) / 2.0; c3 = (3.0 * p->b) / 2.0; c4 = (9.0 * p->c) / 8.0; }
    else { c2 = (3.0 * p->a * dT) / 2.0; c3 = (9.0 * p->b) / 2.0; c4 = (9.0 * p->c) / 4.0; }
    const double i2x = 1.0 / (2.0 * dx), i2y = 1.0 / (2.0 * dy), i2z = 1.0 / (2.0 * dz);
    const size_t sx = (size_t)Ny * Nz, sy = (size_t)Nz;
    const int no_L = (p->L1 == 0.0 && p->L2 == 0.0 && p->L3 == 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 1; i < Nx - 1; ++i)
        for (int j = 1; j < Ny - 1; ++j)
            for (int k = 1; k < Nz - 1; ++k) {
                const size_t id = IDX(i, j, k);
                const double xx = QC(Q, id, 0), xy = QC(Q, id, 1), xz = QC(Q, id, 2);
                const double yy = QC(Q, id, 3), yz = QC(Q, id, 4), zz = -xx - yy;
                const double tr2 = xx * xx + yy * yy + zz * zz + 2 * (xy * xy + xz * xz + yz * yz);
                /* full Q^2 then tr(Q^3) in the order of QSR.cu:904-916 */
                const double q2xx = xx * xx + xy * xy + xz * xz;
                const double q2xy = xx * xy + xy * yy + xz * yz;
                const double q2xz = xx * xz + xy * yz + xz * zz;
                const double q2yx = xy * xx + yy * xy + yz * xz;
                const double q2yy = xy * xy + yy * yy + yz * yz;
                const double q2yz = xy * xz + yy * yz + yz * zz;
                const double q2zx = xz * xx + yz * xy + zz * xz;
                const double q2zy = xz * xy + yz * yy + zz * yz;
                const double q2zz = xz * xz + yz * yz + zz * zz;
                const double tr3 = xx * q2xx + xy * q2yx + xz * q2zx + xy * q2xy + yy * q2yy + yz * q2zy +
                                   xz * q2xz + yz * q2yz + zz * q2zz;
                eb[id] = (c2 * tr2 - c3 * tr3 + c4 * tr2 * tr2) * dV;
                double d[6][3];
                for (int c = 0; c < 5; ++c) {
                    d[c][0] = (QC(Q, id + sx, c) - QC(Q, id - sx, c)) * i2x;
                    d[c][1] = (QC(Q, id + sy, c) - QC(Q, id - sy, c)) * i2y;
                    d[c][2] = (QC(Q, id + 1, c) - QC(Q, id - 1, c)) * i2z;
                }
                for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
#define SQ3(v) ((v)[0] * (v)[0] + (v)[1] * (v)[1] + (v)[2] * (v)[2])
                const double g2 = SQ3(d[0]) + SQ3(d[3]) + SQ3(d[5]) + 2.0 * (SQ3(d[1]) + SQ3(d[2]) + SQ3(d[4]));
#undef SQ3
                double fel;
                if (no_L) {
                    fel = 0.5 * p->kappa * g2;
                } else {
                    const double f1 = 0.5 * p->L1 * g2;
                    const double dvx = d[0][0] + d[1][1] + d[2][2];
                    const double dvy = d[1][0] + d[3][1] + d[4][2];
                    const double dvz = d[2][0] + d[4][1] + d[5][2];
                    const double f2 = 0.5 * p->L2 * (dvx * dvx + dvy * dvy + dvz * dvz);
                    double f3 = 0.0;
                    if (p->L3 != 0.0) {
                        double M[3][3];
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b)
                                M[a][b] = d[0][a] * d[0][b] + d[3][a] * d[3][b] + d[5][a] * d[5][b] +
                                          2.0 * (d[1][a] * d[1][b] + d[2][a] * d[2][b] + d[4][a] * d[4][b]);
                        f3 = 0.5 * p->L3 * (xx * M[0][0] + yy * M[1][1] + zz * M[2][2] +
                                            2.0 * (xy * M[0][1] + xz * M[0][2] + yz * M[1][2]));
                    }
                    fel = f1 + f2 + f3;
                }
                ee[id] = fel * dV;
            }
    if (p->W != 0.0 && thickness > 0.0) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < Nx; ++i)
            for (int j = 0; j < Ny; ++j)
                for (int k = 0; k < Nz; ++k) {
                    const size_t id = IDX(i, j, k);
                    if (!shell[id]) continue;
                    double n[3], q0[5];
                    orc_surface_normal(g, i, j, k, n);
                    uniaxial(S_shell, n, q0);
                    const double q0zz = -(q0[0] + q0[3]);
                    const double a = QC(Q, id, 0) - q0[0], b = QC(Q, id, 1) - q0[1], c = QC(Q, id, 2) - q0[2];
                    const double e = QC(Q, id, 3) - q0[3], f = QC(Q, id, 4) - q0[4];
                    const double zzd = (-(QC(Q, id, 0) + QC(Q, id, 3))) - q0zz;
                    const double diff2 = a * a + e * e + zzd * zzd + 2.0 * (b * b + c * c + f * f);
                    ea[id] = 0.5 * (p->W / thickness) * diff2 * dV;
                }
    }
    double sb = 0.0, se = 0.0, sa = 0.0;
    for (size_t id = 0; id < N; ++id) { sb += eb[id]; se += ee[id]; sa += ea[id]; }
    out[0] = sb; out[1] = se; out[2] = sa;
    free(eb); free(ee); free(ea);
}

/* wrap_to_pi_device (QSR.cu:1027-1035) */
static inline double wrap_floor(double x) {
    const double kPi = 3.141592653589793238462643383279502884, kTwoPi = 2.0 * kPi;
    x = x + kPi;
    x -= floor(x / kTwoPi) * kTwoPi;
    return x - kPi;
}

static inline int nf_ok(const double nf[4], double thr) {
    return isfinite(nf[0]) && isfinite(nf[1]) && isfinite(nf[2]) && (nf[0] > thr);
}

/* computeDefectDensity2DProxyKernel (QSR.cu:1114-1176) */
void orc_defect2d(const double* Q, int Nx, int Ny, int Nz, int z_slice, double S_threshold,
                  double charge_cutoff, unsigned int* defects, unsigned int* used) {
    *defects = 0; *used = 0;
    if (z_slice < 0 || z_slice >= Nz) return;
    const double kInv2Pi = 1.0 / (2.0 * 3.141592653589793238462643383279502884);
    const int k = z_slice;
    unsigned int nd = 0, nu = 0;
#pragma omp parallel for collapse(2) reduction(+ : nd, nu) schedule(static)
    for (int i = 0; i < Nx - 1; ++i)
        for (int j = 0; j < Ny - 1; ++j) {
            double a[4], b[4], c[4], d[4];
            orc_nematic(Q + 5 * IDX(i, j, k), a);
            orc_nematic(Q + 5 * IDX(i + 1, j, k), b);
            orc_nematic(Q + 5 * IDX(i + 1, j + 1, k), c);
            orc_nematic(Q + 5 * IDX(i, j + 1, k), d);
            if (!(nf_ok(a, S_threshold) && nf_ok(b, S_threshold) && nf_ok(c, S_threshold) && nf_ok(d, S_threshold)))
                continue;
            const double p00 = 2.0 * atan2(a[2], a[1]), p10 = 2.0 * atan2(b[2], b[1]);
            const double p11 = 2.0 * atan2(c[2], c[1]), p01 = 2.0 * atan2(d[2], d[1]);
            const double dsum = wrap_floor(p10 - p00) + wrap_floor(p11 - p10) + wrap_floor(p01 - p11) +
                                wrap_floor(p00 - p01);
            const double s = 0.5 * (dsum * kInv2Pi);
            nu += 1u;
            if (fabs(s) > charge_cutoff) nd += 1u;
        }
    *defects = nd; *used = nu;
}

/* computeXiGradEdges{X,Y}Kernel (QSR.cu:1042-1106) + compute_xi_grad_proxy_2d (QSR.cu:2372-2410) */
void orc_xi2d(const double* Q, int Nx, int Ny, int Nz, int z_slice, double S_threshold,
              double* xi, unsigned int* edges) {
    *xi = NAN; *edges = 0;
    if (z_slice < 0 || z_slice >= Nz) return;
    const int k = z_slice;
    double sum = 0.0;
    unsigned int ne = 0;
    for (int dir = 0; dir < 2; ++dir) {
        const int imax = (dir == 0) ? Nx - 1 : Nx;
        const int jmax = (dir == 0) ? Ny : Ny - 1;
        for (int i = 0; i < imax; ++i)
            for (int j = 0; j < jmax; ++j) {
                double a[4], b[4];
                orc_nematic(Q + 5 * IDX(i, j, k), a);
                if (dir == 0) orc_nematic(Q + 5 * IDX(i + 1, j, k), b);
                else orc_nematic(Q + 5 * IDX(i, j + 1, k), b);
                if (!(nf_ok(a, S_threshold) && nf_ok(b, S_threshold))) continue;
                const double dd = wrap_floor(2.0 * atan2(b[2], b[1]) - 2.0 * atan2(a[2], a[1]));
                sum += dd * dd;
                ne += 1u;
            }
    }
    *edges = ne;
    if (ne > 0u) {
        const double mean_sq = sum / (double)ne;
        const double eps = 1e-30;
        const double inv = (mean_sq > 0.0) ? (1.0 / sqrt(mean_sq + eps)) : (1.0 / sqrt(eps));
        const double cap = 0.5 * (double)(Nx < Ny ? Nx : Ny);
        *xi = inv < cap ? inv : cap;
    }
}

/* computeRadialityKernel (QSR.cu:994-1015) + host average (QSR.cu:3246-3256) */
double orc_radiality(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g) {
    double tot = 0.0;
    long long cnt = 0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                double s[5];
                orc_sample(g, i, j, k, s);
                if (!(s[4] <= 0.0 && s[3] > 2.0 * g->min_spacing)) continue;
                double nf[4], r[3];
                orc_nematic(Q + 5 * IDX(i, j, k), nf);
                center_direction(s, r);
                tot += fabs(nf[1] * r[0] + nf[2] * r[1] + nf[3] * r[2]);
                cnt += 1;
            }
    return (cnt > 0) ? tot / (double)cnt : 0.0;
}

/* compute_avg_S_droplet / compute_max_S_droplet (QSR.cu:2285-2321): out = {avg, max} */
void orc_avg_max_S(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, double out[2]) {
    const double excl = 2.0 * g->min_spacing;
    double tot = 0.0, mx = 0.0;
    long long cnt = 0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                double s[5];
                orc_sample(g, i, j, k, s);
                if (!(s[4] < -excl)) continue;
                double nf[4];
                orc_nematic(Q + 5 * IDX(i, j, k), nf);
                tot += nf[0];
                cnt += 1;
                if (isfinite(nf[0]) && nf[0] > mx) mx = nf[0];
            }
    out[0] = (cnt > 0) ? tot / (double)cnt : 0.0;
    out[1] = mx;
}

/* checkerboard metric (QSR.cu:2846-2873) */
double orc_checker(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, int z_check,
                   double max_S) {
    if (!(max_S >= 0.5)) return NAN;
    if (z_check < 0 || z_check >= Nz) return NAN;
    const double excl = 2.0 * g->min_spacing;
    double es = 0.0, os = 0.0, as = 0.0;
    int en = 0, on = 0, an = 0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j) {
            double s[5];
            orc_sample(g, i, j, z_check, s);
            if (s[4] >= -excl) continue;
            double nf[4];
            orc_nematic(Q + 5 * IDX(i, j, z_check), nf);
            if (!isfinite(nf[0]) || nf[0] <= 0.01) continue;
            as += nf[0]; an++;
            if (((i + j) & 1) == 0) { es += nf[0]; en++; }
            else { os += nf[0]; on++; }
        }
    if (en > 10 && on > 10 && an > 20 && as != 0.0) return fabs(es - os) / fabs(as);
    return NAN;
}

/* ------------------------------------------------------------------------- */
/* Seeded initial states: QSR.cu:1834-1904 with the libstdc++ algorithms for  */
/* std::mt19937, std::generate_canonical<double,53>,                          */
/* std::uniform_real_distribution<double> and std::normal_distribution<double>*/
/* (Marsaglia polar method with one cached value).                           */
/* ------------------------------------------------------------------------- */

typedef struct { uint32_t mt[624]; int idx; } orc_mt;

/* MT19937 (Matsumoto & Nishimura 1998), seeding as in [rand.eng.mers] */
static void mt_seed(orc_mt* m, uint32_t s) {
    m->mt[0] = s;
    for (int i = 1; i < 624; ++i) m->mt[i] = 1812433253u * (m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) + (uint32_t)i;
    m->idx = 624;
}

static uint32_t mt_next(orc_mt* m) {
    if (m->idx >= 624) {
        for (int i = 0; i < 624; ++i) {
            const uint32_t y = (m->mt[i] & 0x80000000u) | (m->mt[(i + 1) % 624] & 0x7fffffffu);
            m->mt[i] = m->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        m->idx = 0;
    }
    uint32_t y = m->mt[m->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* generate_canonical<double, 53> over a 32-bit engine: two draws, k = 2 */
static double mt_canonical(orc_mt* m) {
    double sum = 0.0, tmp = 1.0;
    for (int t = 0; t < 2; ++t) {
        sum += (double)mt_next(m) * tmp;
        tmp *= 4294967296.0;
    }
    double r = sum / tmp;
    if (r >= 1.0) r = nextafter(1.0, 0.0);
    return r;
}

typedef struct { orc_mt mt; int have_saved; double saved; } orc_normal_gen;

static double normal_draw(orc_normal_gen* g, double mean, double stddev) {
    double ret;
    if (g->have_saved) {
        g->have_saved = 0;
        ret = g->saved;
    } else {
        double x, y, r2;
        do {
            x = 2.0 * mt_canonical(&g->mt) - 1.0;
            y = 2.0 * mt_canonical(&g->mt) - 1.0;
            r2 = x * x + y * y;
        } while (r2 > 1.0 || r2 == 0.0);
        const double mult = sqrt(-2.0 * log(r2) / r2);
        g->saved = x * mult;
        g->have_saved = 1;
        ret = y * mult;
    }
    return ret * stddev + mean;
}

void orc_mt19937_stream(uint32_t seed, uint32_t* out, size_t n) {
    orc_mt m;
    mt_seed(&m, seed);
    for (size_t t = 0; t < n; ++t) out[t] = mt_next(&m);
}

void orc_canonical_stream(uint32_t seed, double* out, size_t n) {
    orc_mt m;
    mt_seed(&m, seed);
    for (size_t t = 0; t < n; ++t) out[t] = mt_canonical(&m);
}

void orc_normal_stream(uint32_t seed, double mean, double stddev, double* out, size_t n) {
    orc_normal_gen g;
    mt_seed(&g.mt, seed);
    g.have_saved = 0; g.saved = 0.0;
    for (size_t t = 0; t < n; ++t) out[t] = normal_draw(&g, mean, stddev);
}

/* initializeQTensor (QSR.cu:1841-1878); seed >= 0 required (QSR.cu:1834-1839) */
void orc_init_director(double* Q, int Nx, int Ny, int Nz, double q0, int mode, const orc_geom* g,
                       long long seed) {
    const size_t N = (size_t)Nx * Ny * Nz;
    memset(Q, 0, 5 * N * sizeof(double));
    orc_mt m;
    mt_seed(&m, (uint32_t)seed);
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                double s[5];
                orc_sample(g, i, j, k, s);
                if (!(s[4] <= 0.0)) continue;
                double S = q0;
                if (s[3] < g->core_radius) S *= (s[3] / g->core_radius);
                double n[3];
                if (mode == 0) {
                    const double theta = mt_canonical(&m) * (ORC_PI - 0.0) + 0.0;
                    const double phi = mt_canonical(&m) * (2.0 * ORC_PI - 0.0) + 0.0;
                    n[0] = sin(theta) * cos(phi);
                    n[1] = sin(theta) * sin(phi);
                    n[2] = cos(theta);
                } else {
                    center_direction(s, n);
                }
                uniaxial(S, n, Q + 5 * IDX(i, j, k));
            }
}

/* initializeIsotropicWithNoise (QSR.cu:1880-1904) */
void orc_init_isotropic(double* Q, int Nx, int Ny, int Nz, double amp, const orc_geom* g,
                        long long seed) {
    const size_t N = (size_t)Nx * Ny * Nz;
    memset(Q, 0, 5 * N * sizeof(double));
    if (amp <= 0.0) return;
    orc_normal_gen gen;
    mt_seed(&gen.mt, (uint32_t)seed);
    gen.have_saved = 0; gen.saved = 0.0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                double s[5];
                orc_sample(g, i, j, k, s);
                if (!(s[4] <= 0.0)) continue;
                double* q = Q + 5 * IDX(i, j, k);
                for (int c = 0; c < 5; ++c) q[c] = normal_draw(&gen, 0.0, amp);
            }
}

/* ------------------------------------------------------------------------- */
/* Periodic bulk branch (KZM_bulk_ldg.cu)                                     */
/* ------------------------------------------------------------------------- */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Five N(0,1) variates for (seed, voxel, step): two Philox blocks with counter
 * {voxel_lo, voxel_hi, step, block}, key {seed_lo, seed_hi}; uniforms
 * u = ((x >> 9) + 0.5) * 2^-23 in (0,1) (exactly representable in binary32, so
 * an FP32 device implementation sees the same uniforms); Box-Muller pairs
 * (u0,u1), (u2,u3), (u4,u5). */
void orc_noise5(uint64_t seed, uint64_t voxel, uint32_t step, double g[5]) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)voxel, (uint32_t)(voxel >> 32), step, 0u};
    uint32_t r[4];
    orc_philox4x32_10(ctr, key, r);
    /* the 128-bit block as r0 + 2^32 r1 + 2^64 r2 + 2^96 r3, cut into six 21-bit fields */
    double u[6];
    for (int t = 0; t < 6; ++t) {
        const int bit = 21 * t, w = bit / 32, sh = bit % 32;
        uint64_t two = r[w];
        if (w + 1 < 4) two |= (uint64_t)r[w + 1] << 32;
        const uint32_t f = (uint32_t)(two >> sh) & 0x1FFFFFu;
        u[t] = ((double)f + 0.5) * (1.0 / 2097152.0);
    }
    for (int pr = 0; pr < 3; ++pr) {
        const double rad = sqrt(-2.0 * log(u[2 * pr]));
        const double ang = 2.0 * ORC_PI * u[2 * pr + 1];
        const double g0 = rad * cos(ang), g1 = rad * sin(ang);
        if (2 * pr < 5) g[2 * pr] = g0;
        if (2 * pr + 1 < 5) g[2 * pr + 1] = g1;
    }
}

/* updateBulkLdgKernel (KZM_bulk_ldg.cu:351-469), periodic wrap on all faces.
 * Noise is Philox-keyed (orc_noise5) instead of the reference's splitmix64. */
void orc_periodic_step(const double* cur, double* next, int Nx, int Ny, int Nz,
                       double dx, double dy, double dz, double A, double B, double C, double kappa,
                       double gamma, double dt, double noise_sigma, uint64_t seed, uint32_t step,
                       double q_norm_cap) {
    const double ix2 = 1.0 / (dx * dx), iy2 = 1.0 / (dy * dy), iz2 = 1.0 / (dz * dz);
    const double mob = 1.0 / gamma;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                const int ip = (i + 1) % Nx, im = (i + Nx - 1) % Nx;
                const int jp = (j + 1) % Ny, jm = (j + Ny - 1) % Ny;
                const int kp = (k + 1) % Nz, km = (k + Nz - 1) % Nz;
                const double* c0 = cur + 5 * id;
                double lap[5];
                for (int c = 0; c < 5; ++c)
                    lap[c] = (QC(cur, IDX(ip, j, k), c) - 2.0 * c0[c] + QC(cur, IDX(im, j, k), c)) * ix2 +
                             (QC(cur, IDX(i, jp, k), c) - 2.0 * c0[c] + QC(cur, IDX(i, jm, k), c)) * iy2 +
                             (QC(cur, IDX(i, j, kp), c) - 2.0 * c0[c] + QC(cur, IDX(i, j, km), c)) * iz2;
                const double xx = c0[0], xy = c0[1], xz = c0[2], yy = c0[3], yz = c0[4], zz = -(xx + yy);
                const double q2xx = xx * xx + xy * xy + xz * xz;
                const double q2xy = xx * xy + xy * yy + xz * yz;
                const double q2xz = xx * xz + xy * yz + xz * zz;
                const double q2yy = xy * xy + yy * yy + yz * yz;
                const double q2yz = xy * xz + yy * yz + yz * zz;
                const double q2zz = xz * xz + yz * yz + zz * zz;
                const double tr2 = xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz);
                double hxx = A * xx - B * (q2xx - tr2 / 3.0) + C * tr2 * xx - kappa * lap[0];
                double hyy = A * yy - B * (q2yy - tr2 / 3.0) + C * tr2 * yy - kappa * lap[3];
                const double hzz = A * zz - B * (q2zz - tr2 / 3.0) + C * tr2 * zz + kappa * (lap[0] + lap[3]);
                const double hxy = A * xy - B * q2xy + C * tr2 * xy - kappa * lap[1];
                const double hxz = A * xz - B * q2xz + C * tr2 * xz - kappa * lap[2];
                const double hyz = A * yz - B * q2yz + C * tr2 * yz - kappa * lap[4];
                const double corr = (hxx + hyy + hzz) / 3.0;
                hxx -= corr; hyy -= corr;
                double* o = next + 5 * id;
                o[0] = xx - mob * dt * hxx;
                o[1] = xy - mob * dt * hxy;
                o[2] = xz - mob * dt * hxz;
                o[3] = yy - mob * dt * hyy;
                o[4] = yz - mob * dt * hyz;
                if (noise_sigma > 0.0) {
                    double gg[5];
                    orc_noise5(seed, (uint64_t)id, step, gg);
                    for (int c = 0; c < 5; ++c) o[c] += noise_sigma * gg[c];
                }
                if (q_norm_cap > 0.0) {
                    const double nrm = frob(o);
                    if (nrm > q_norm_cap && nrm > 0.0) {
                        const double sc = q_norm_cap / nrm;
                        for (int c = 0; c < 5; ++c) o[c] *= sc;
                    }
                }
            }
}

/* calculateNematicFieldDevice variant of KZM_bulk_ldg.cu:200-252 (isotropic
 * threshold 1e-12, isotropic result S=0, n=(1,0,0)) */
static void nematic_kzm(const double* q5, double out[4]) {
    if (fabs(q5[0]) < 1e-12 && fabs(q5[1]) < 1e-12 && fabs(q5[2]) < 1e-12 && fabs(q5[3]) < 1e-12 &&
        fabs(q5[4]) < 1e-12) {
        out[0] = 0.0; out[1] = 1.0; out[2] = 0.0; out[3] = 0.0;
        return;
    }
    /* for |Q| >= 1e-9 the algorithm equals QSR.cu:191-243; between 1e-12 and
     * 1e-9 QSR.cu short-cuts, so run the iteration explicitly here */
    const double xx = q5[0], xy = q5[1], xz = q5[2], yy = q5[3], yz = q5[4], zz = -xx - yy;
    double nx = xx + xy + xz, ny = xy + yy + yz, nz = xz + yz - xx - yy;
    const double n0 = sqrt(nx * nx + ny * ny + nz * nz);
    if (n0 > 1e-12) { nx /= n0; ny /= n0; nz /= n0; }
    else { nx = 0.0; ny = 0.0; nz = 1.0; }
    const double r1 = fabs(xx) + fabs(xy) + fabs(xz);
    const double r2 = fabs(xy) + fabs(yy) + fabs(yz);
    const double r3 = fabs(xz) + fabs(yz) + fabs(zz);
    const double alpha = fmax(r1, fmax(r2, r3)) + 1e-12;
    for (int it = 0; it < 15; ++it) {
        const double ax = xx * nx + xy * ny + xz * nz + alpha * nx;
        const double ay = xy * nx + yy * ny + yz * nz + alpha * ny;
        const double az = xz * nx + yz * ny + zz * nz + alpha * nz;
        const double nrm = sqrt(ax * ax + ay * ay + az * az);
        if (nrm < 1e-13) break;
        nx = ax / nrm; ny = ay / nrm; nz = az / nrm;
    }
    const double lam = nx * (xx * nx + xy * ny + xz * nz) + ny * (xy * nx + yy * ny + yz * nz) +
                       nz * (xz * nx + yz * ny + zz * nz);
    if (nz < -1e-10) { nx = -nx; ny = -ny; nz = -nz; }
    out[0] = 1.5 * lam; out[1] = nx; out[2] = ny; out[3] = nz;
}

/* wrap_to_pi (KZM_bulk_ldg.cu:471-479) */
static inline double wrap_loop(double d) {
    while (d <= -ORC_PI) d += 2.0 * ORC_PI;
    while (d > ORC_PI) d -= 2.0 * ORC_PI;
    return d;
}

/* project_to_plane + plaquette_charge (KZM_bulk_ldg.cu:481-541).
 * returns -1 unusable, 0 no defect, 1 defect */
static int plaquette(const double* a, const double* b, const double* c, const double* d, int plane,
                     double thr, double cutoff) {
    const double* nf[4] = {a, b, c, d};
    for (int t = 0; t < 4; ++t)
        if (!isfinite(nf[t][0])) return -1;
    for (int t = 0; t < 4; ++t)
        if (!(nf[t][0] > thr)) return -1;
    double p[4];
    for (int t = 0; t < 4; ++t) {
        double u, v;
        if (plane == 0) { u = nf[t][1]; v = nf[t][2]; }
        else if (plane == 1) { u = nf[t][2]; v = nf[t][3]; }
        else { u = nf[t][1]; v = nf[t][3]; }
        const double nrm = sqrt(u * u + v * v);
        if (!(nrm > 1e-8)) return -1;
        p[t] = 2.0 * atan2(v / nrm, u / nrm);
    }
    const double dsum = wrap_loop(p[1] - p[0]) + wrap_loop(p[2] - p[1]) + wrap_loop(p[3] - p[2]) +
                        wrap_loop(p[0] - p[3]);
    const double s = 0.5 * (dsum / (2.0 * ORC_PI));
    return fabs(s) > cutoff ? 1 : 0;
}

/* compute_observables (KZM_bulk_ldg.cu:543-713), x <-> i, y <-> j, z <-> k */
void orc_periodic_observables(const double* Q, int Nx, int Ny, int Nz, double dx, double dy,
                              double dz, double A, double B, double C, double kappa,
                              double S_threshold, double charge_cutoff, double out[10]) {
    const size_t N = (size_t)Nx * Ny * Nz;
    const double dV = dx * dy * dz, V = dV * (double)Nx * (double)Ny * (double)Nz;
    const double ix = 1.0 / dx, iy = 1.0 / dy, iz = 1.0 / dz;
    double* nf = (double*)malloc(4 * N * sizeof(double));
    double bulk = 0.0, el = 0.0, qn = 0.0, gs = 0.0, Ssum = 0.0, Smax = 0.0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const size_t id = IDX(i, j, k);
                const double* c0 = Q + 5 * id;
                const double* px = Q + 5 * IDX((i + 1) % Nx, j, k);
                const double* py = Q + 5 * IDX(i, (j + 1) % Ny, k);
                const double* pz = Q + 5 * IDX(i, j, (k + 1) % Nz);
                const double xx = c0[0], xy = c0[1], xz = c0[2], yy = c0[3], yz = c0[4], zz = -(xx + yy);
                const double q2xx = xx * xx + xy * xy + xz * xz;
                const double q2xy = xx * xy + xy * yy + xz * yz;
                const double q2xz = xx * xz + xy * yz + xz * zz;
                const double q2yy = xy * xy + yy * yy + yz * yz;
                const double q2yz = xy * xz + yy * yz + yz * zz;
                const double q2zz = xz * xz + yz * yz + zz * zz;
                const double tr2 = xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz);
                const double tr3 = xx * q2xx + xy * q2xy + xz * q2xz + xy * q2xy + yy * q2yy + yz * q2yz +
                                   xz * q2xz + yz * q2yz + zz * q2zz;
                bulk += ((A / 2.0) * tr2 - (B / 3.0) * tr3 + (C / 4.0) * tr2 * tr2) * dV;
                double d[6][3];
                for (int c = 0; c < 5; ++c) {
                    d[c][0] = (px[c] - c0[c]) * ix;
                    d[c][1] = (py[c] - c0[c]) * iy;
                    d[c][2] = (pz[c] - c0[c]) * iz;
                }
                for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
                double g2 = 0.0;
                for (int a = 0; a < 3; ++a)
                    g2 += d[0][a] * d[0][a] + d[3][a] * d[3][a] + d[5][a] * d[5][a];
                double g2o = 0.0;
                for (int a = 0; a < 3; ++a)
                    g2o += d[1][a] * d[1][a] + d[2][a] * d[2][a] + d[4][a] * d[4][a];
                g2 += 2.0 * g2o;
                gs += g2 * dV;
                el += 0.5 * kappa * g2 * dV;
                qn += tr2 * dV;
                nematic_kzm(c0, nf + 4 * id);
                Ssum += nf[4 * id];
                if (nf[4 * id] > Smax) Smax = nf[4 * id];
            }
    long long defects = 0, used = 0;
    double line = 0.0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const int ip = (i + 1) % Nx, jp = (j + 1) % Ny, kp = (k + 1) % Nz;
                const double* n000 = nf + 4 * IDX(i, j, k);
                const double* n100 = nf + 4 * IDX(ip, j, k);
                const double* n110 = nf + 4 * IDX(ip, jp, k);
                const double* n010 = nf + 4 * IDX(i, jp, k);
                const double* n001 = nf + 4 * IDX(i, j, kp);
                const double* n101 = nf + 4 * IDX(ip, j, kp);
                const double* n011 = nf + 4 * IDX(i, jp, kp);
                int r = plaquette(n000, n100, n110, n010, 0, S_threshold, charge_cutoff);
                if (r >= 0) { used++; if (r) { defects++; line += dz; } }
                r = plaquette(n000, n010, n011, n001, 1, S_threshold, charge_cutoff);
                if (r >= 0) { used++; if (r) { defects++; line += dx; } }
                r = plaquette(n000, n100, n101, n001, 2, S_threshold, charge_cutoff);
                if (r >= 0) { used++; if (r) { defects++; line += dy; } }
            }
    free(nf);
    out[0] = bulk / V;
    out[1] = el / V;
    out[2] = out[0] + out[1];
    out[3] = Ssum / (double)N;
    out[4] = Smax;
    out[5] = (used > 0) ? (double)defects / (double)used : NAN;
    out[6] = (double)used;
    out[7] = (double)defects;
    out[8] = (gs > 0.0 && qn > 0.0) ? sqrt(qn / gs) : NAN;
    out[9] = (V > 0.0) ? line / V : NAN;
}
