// ref_harness.cpp -- C-ABI shim over the reference's own CPU functions in
// QSR.cpp (compiled from /root/reference by oracle/build_ref.sh, with QSR.cpp's
// interactive main() renamed).  TEST INFRASTRUCTURE ONLY: it produces golden
// vectors for the shared arithmetic (QSR.cpp:10-61, 179-282, 411-557, 559-826)
// and times the reference CPU step for the CPU baseline.  Nothing in the product
// links it.
#include "QSR.h"

#include <omp.h>

#include <cstring>
#include <vector>

static std::vector<QTensor> to_vec(const double* q, size_t n) {
    std::vector<QTensor> v(n);
    for (size_t t = 0; t < n; ++t) {
        v[t].Qxx = q[5 * t + 0];
        v[t].Qxy = q[5 * t + 1];
        v[t].Qxz = q[5 * t + 2];
        v[t].Qyy = q[5 * t + 3];
        v[t].Qyz = q[5 * t + 4];
    }
    return v;
}

static void from_vec(const std::vector<QTensor>& v, double* q) {
    for (size_t t = 0; t < v.size(); ++t) {
        q[5 * t + 0] = v[t].Qxx;
        q[5 * t + 1] = v[t].Qxy;
        q[5 * t + 2] = v[t].Qxz;
        q[5 * t + 3] = v[t].Qyy;
        q[5 * t + 4] = v[t].Qyz;
    }
}

static DimensionalParams make_params(const double* p) {
    // p = {a, b, c, T, T_star, L1, L2, L3, W}
    DimensionalParams d;
    d.a = p[0]; d.b = p[1]; d.c = p[2]; d.T = p[3]; d.T_star = p[4];
    d.L1 = p[5]; d.L2 = p[6]; d.L3 = p[7]; d.W = p[8];
    return d;
}

extern "C" {

int ref_max_threads() { return omp_get_max_threads(); }

void ref_laplacian(const double* Q, double* lap, int Nx, int Ny, int Nz, double dx, double dy, double dz) {
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<QTensor> q = to_vec(Q, n), l(n);
    computeLaplacian(q, l, Nx, Ny, Nz, dx, dy, dz);
    from_vec(l, lap);
}

void ref_chempot(const double* Q, const double* lap, double* mu, int Nx, int Ny, int Nz, double dx,
                 double dy, double dz, const double* params9, double kappa, int modechoice) {
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<QTensor> q = to_vec(Q, n), l = to_vec(lap, n), m(n);
    computeChemicalPotential(q, m, l, Nx, Ny, Nz, dx, dy, dz, make_params(params9), kappa, modechoice);
    from_vec(m, mu);
}

void ref_energy(const double* Q, int Nx, int Ny, int Nz, double dx, double dy, double dz,
                const double* params9, double kappa, int modechoice, double* out2) {
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<QTensor> q = to_vec(Q, n);
    EnergyComponents e = computeEnergyComponents(q, Nx, Ny, Nz, dx, dy, dz, make_params(params9), kappa, modechoice);
    out2[0] = e.bulk;
    out2[1] = e.elastic;
}

void ref_nematic(const double* Q, size_t n, double* out4) {
    for (size_t t = 0; t < n; ++t) {
        QTensor q;
        q.Qxx = Q[5 * t + 0]; q.Qxy = Q[5 * t + 1]; q.Qxz = Q[5 * t + 2];
        q.Qyy = Q[5 * t + 3]; q.Qyz = Q[5 * t + 4];
        NematicField nf = calculateNematicField(q);
        out4[4 * t + 0] = nf.S; out4[4 * t + 1] = nf.nx; out4[4 * t + 2] = nf.ny; out4[4 * t + 3] = nf.nz;
    }
}

// QSR.cpp:861-891 updateQTensor on the given field, in place.  shell is the
// caller's mask in the reference index order (1 = shell voxel).
void ref_update(double* Q, const double* mu, int Nx, int Ny, int Nz, double dt, double gamma, double W,
                const unsigned char* shell) {
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<QTensor> q = to_vec(Q, n), m = to_vec(mu, n);
    std::vector<bool> sh(n);
    for (size_t t = 0; t < n; ++t) sh[t] = shell[t] != 0;
    updateQTensor(q, m, Nx, Ny, Nz, dt, sh, gamma, W);
    from_vec(q, Q);
}

// Time `steps` reference CPU iterations (QSR.cpp loop body: laplacian, chemical
// potential incl. L2/L3, weak anchoring, update, strong BC) on the given field.
// The field is advanced in place.  Returns wall seconds.
double ref_time_steps(double* Q, int Nx, int Ny, int Nz, double dx, double dy, double dz,
                      const double* params9, double kappa, int modechoice, double dt, double gamma,
                      double S_shell, int steps) {
    const size_t n = (size_t)Nx * Ny * Nz;
    std::vector<QTensor> q = to_vec(Q, n), lap(n), mu(n);
    const DimensionalParams p = make_params(params9);
    // QSR.cpp shell mask convention (QSR.cpp:1091-1102): |r - R| < 1 about N/2
    std::vector<bool> shell(n, false);
    const double R = std::min({(double)Nx, (double)Ny, (double)Nz}) / 2.0;
    for (int i = 0; i < Nx; ++i)
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const double x = i - Nx / 2.0, y = j - Ny / 2.0, z = k - Nz / 2.0;
                const double r = std::sqrt(x * x + y * y + z * z);
                shell[(size_t)k + (size_t)Nz * ((size_t)j + (size_t)Ny * i)] = std::fabs(r - R) < 1.0;
            }
    const double t0 = omp_get_wtime();
    for (int s = 0; s < steps; ++s) {
        computeLaplacian(q, lap, Nx, Ny, Nz, dx, dy, dz);
        computeChemicalPotential(q, mu, lap, Nx, Ny, Nz, dx, dy, dz, p, kappa, modechoice);
        applyWeakAnchoringPenalty(mu, q, Nx, Ny, Nz, shell, S_shell, p.W);
        updateQTensor(q, mu, Nx, Ny, Nz, dt, shell, gamma, p.W);
        applyBoundaryConditions(q, Nx, Ny, Nz, shell, S_shell, p.W);
    }
    const double t1 = omp_get_wtime();
    from_vec(q, Q);
    return t1 - t0;
}

}  // extern "C"
