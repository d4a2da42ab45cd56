/*
 * qsr_oracle.h -- FP64 CPU oracle for the QSR Landau-de Gennes relaxation /
 * quench step.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP product path is compared against (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  Nothing in the product links or calls it.
 *
 * Storage convention (the reference's interchange format, QSR.cuh:22-44):
 *   Q is an array of 5-double records {Qxx, Qxy, Qxz, Qyy, Qyz} (AoS),
 *   linear index idx = k + Nz*(j + Ny*i)  (QSR.cu:286), z fastest.
 * Every function names the reference file:line it restates.
 */
#ifndef QSR_ORACLE_H
#define QSR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double dx, dy, dz;
    double center_x, center_y, center_z;
    double radius_x, radius_y, radius_z;
    double min_spacing, min_radius, shell_half_thickness, core_radius;
} orc_geom; /* QSR.cu:34-48 */

typedef struct {
    double a, b, c, T, T_star;
    int modechoice;
    double kappa, L1, L2, L3, W;
} orc_phys; /* QSR.cuh:84-96 + kappa/modechoice kernel arguments */

/* ---- geometry (QSR.cu:70-185) ---- */
void orc_make_geometry(int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double radius_index, double cx_index, double cy_index, double cz_index,
                       orc_geom* g);
void orc_sample(const orc_geom* g, int i, int j, int k, double out[5]); /* x,y,z,r,sd */
int orc_surface_normal(const orc_geom* g, int i, int j, int k, double n[3]);
size_t orc_shell_mask(const orc_geom* g, int Nx, int Ny, int Nz, uint8_t* shell);
double orc_surface_area(const orc_geom* g);
double orc_shell_thickness(const orc_geom* g, size_t shell_count);

/* ---- bulk physics host helpers (QSR.cu:1363-1604) ---- */
void orc_bulk_ABC(double a, double b, double c, double T, double T_star, int mode, double out[3]);
double orc_bulk_fS(double A, double B, double C, double S);
double orc_S_eq(double A, double B, double C);
double orc_A_NI(double B, double C);
double orc_A_spinN(double B, double C);
double orc_T_from_A(double A, double a, double T_star);
double orc_shell_order(int custom, double custom_S, double a, double b, double c,
                       double T, double T_star, int mode);
/* out: [min_dx, S_scale, L_eff_total, D_total, dt_diff_total, L_eff_explicit,
 *       dt_diff, bulk_rate, dt_bulk, dt_max] */
void orc_dt_estimate(double dx, double dy, double dz, double gamma, double a, double b, double c,
                     double T_star, int mode, double kappa, double L1, double L2, double L3,
                     double S_ref_used, double S0, double T_a, double T_b,
                     int use_semi_implicit, double L_stab, int jacobi_iters, double out[10]);

/* ---- step kernels, one pass each (QSR.cu:279-867) ---- */
void orc_laplacian(const double* Q, double* lap, int Nx, int Ny, int Nz, double dx, double dy, double dz);
void orc_divergence(const double* Q, double* Dx, double* Dy, double* Dz,
                    int Nx, int Ny, int Nz, double dx, double dy, double dz);
void orc_chempot(const double* Q, double* mu, const double* lap, int Nx, int Ny, int Nz,
                 double dx, double dy, double dz, const orc_phys* p,
                 const double* Dx, const double* Dy, const double* Dz);
void orc_chempot_l3(const double* Q, double* mu, int Nx, int Ny, int Nz,
                    double dx, double dy, double dz, const orc_phys* p,
                    const double* Dx, const double* Dy, const double* Dz);
void orc_weak_anchor(double* mu, const double* Q, int Nx, int Ny, int Nz, const uint8_t* shell,
                     const orc_geom* g, double S_shell, double W, double thickness);
void orc_update(double* Q, const double* mu, int Nx, int Ny, int Nz, double dt, const uint8_t* shell,
                const orc_geom* g, double gamma, double W, double q_norm_cap);
void orc_bc(double* Q, int Nx, int Ny, int Nz, const uint8_t* shell, const orc_geom* g,
            double S_shell, double W);
void orc_rhs_semi_implicit(const double* Q, const double* mu, const double* lap, double* rhs,
                           int Nx, int Ny, int Nz, double dt, double gamma, double L_stab,
                           const uint8_t* shell, const orc_geom* g, double W);
void orc_jacobi_step(const double* rhs, const double* Qold, double* Qnew, int Nx, int Ny, int Nz,
                     double alpha, double dx, double dy, double dz, const uint8_t* shell,
                     const orc_geom* g, double W);

/* One full explicit-Euler iteration in the reference launch order
 * (QSR.cu:3223-3238).  work must hold 13*N doubles (lap, mu, Dcol x3). */
void orc_step(double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
              const uint8_t* shell, double thickness, double S_shell, double dt, double gamma,
              double q_norm_cap, double* work);
/* One IMEX iteration (QSR.cu:2713-2753): work must hold 23*N doubles. */
void orc_step_imex(double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
                   const uint8_t* shell, double thickness, double S_shell, double dt, double gamma,
                   double L_stab, int jacobi_iters, double* work);

/* ---- observables (QSR.cu:191-243, 654-695, 869-1176, 2285-2410, 2829-2874) ---- */
void orc_nematic(const double q5[5], double out[4]);
/* out: bulk, elastic, anchoring (sequential host sums, QSR.cu:2323-2341) */
void orc_energy(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, const orc_phys* p,
                const uint8_t* shell, double S_shell, double thickness, double out[3]);
void orc_defect2d(const double* Q, int Nx, int Ny, int Nz, int z_slice, double S_threshold,
                  double charge_cutoff, unsigned int* defects, unsigned int* used);
void orc_xi2d(const double* Q, int Nx, int Ny, int Nz, int z_slice, double S_threshold,
              double* xi, unsigned int* edges);
double orc_radiality(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g);
void orc_avg_max_S(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, double out[2]);
double orc_checker(const double* Q, int Nx, int Ny, int Nz, const orc_geom* g, int z_check,
                   double max_S);

/* ---- seeded initial states (QSR.cu:1834-1904) with libstdc++ draw order ---- */
void orc_init_director(double* Q, int Nx, int Ny, int Nz, double q0, int mode, const orc_geom* g,
                       long long seed);
void orc_init_isotropic(double* Q, int Nx, int Ny, int Nz, double amp, const orc_geom* g,
                        long long seed);
/* raw generators exposed for pinning */
void orc_mt19937_stream(uint32_t seed, uint32_t* out, size_t n);
void orc_canonical_stream(uint32_t seed, double* out, size_t n);
void orc_normal_stream(uint32_t seed, double mean, double stddev, double* out, size_t n);

/* ---- periodic bulk branch (KZM_bulk_ldg.cu:351-469, 543-713) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Five N(0,1) variates for (seed, voxel, step): one Philox4x32-10 block, counter
 * (voxel lo, voxel hi, step, 0), key (seed lo, seed hi); six 21-bit fields of the
 * 128 bits -> uniforms (f + 1/2) 2^-21 -> three Box-Muller pairs.  This build's
 * noise design (the reference's periodic benchmark uses splitmix64). */
void orc_noise5(uint64_t seed, uint64_t voxel, uint32_t step, double g[5]);
void orc_periodic_step(const double* cur, double* next, int Nx, int Ny, int Nz,
                       double dx, double dy, double dz, double A, double B, double C, double kappa,
                       double gamma, double dt, double noise_sigma, uint64_t seed, uint32_t step,
                       double q_norm_cap);
/* out: bulk/V, elastic/V, total/V, avg_S, max_S, defect_density, plaquettes_used,
 *      defects, xi_grad_proxy, defect_line_density */
void orc_periodic_observables(const double* Q, int Nx, int Ny, int Nz, double dx, double dy,
                              double dz, double A, double B, double C, double kappa,
                              double S_threshold, double charge_cutoff, double out[10]);

#ifdef __cplusplus
}
#endif
#endif
