// qsr_host.hpp -- host-side physics and geometry of the QSR LdG engine:
// bulk constants and S_eq, transition temperatures, shell-order policy, the
// explicit-step stability estimate, the droplet geometry with its mask encoded
// as per-row k-intervals, and the seeded initial states.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace qsrh {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// Bulk Landau-de Gennes constants (QSR.cu:1363-1487)
// ---------------------------------------------------------------------------
struct BulkABC {
    double A, B, C;
};
// f = A/2 trQ^2 - B/3 trQ^3 + C/4 (trQ^2)^2 with A = 3a(T-T*), B/C per convention.
BulkABC bulk_abc(double a, double b, double c, double T, double T_star, int modechoice);
double bulk_f_uniaxial(const BulkABC& k, double S);        // f(S) for Q = S(nn - I/3)
double s_eq(const BulkABC& k);                             // global-minimum uniaxial S, else 0
double a_coexistence(double B, double C);                  // A at T_NI (B^2 / 27C)
double a_spinodal_nematic(double B, double C);             // A at T_spin,N (B^2 / 24C)
double t_from_a(double A, double a, double T_star);
double shell_order(bool custom, double custom_S, double a, double b, double c, double T,
                   double T_star, int modechoice);

// ---------------------------------------------------------------------------
// Explicit stability window (QSR.cu:1489-1604)
// ---------------------------------------------------------------------------
struct DtInput {
    double dx = 1, dy = 1, dz = 1, gamma = 1;
    double a = 0, b = 0, c = 0, T_star = 0;
    int modechoice = 1;
    double kappa = 0, L1 = 0, L2 = 0, L3 = 0;
    double S_ref_used = 0;  // NaN when no Frank mapping
    double S0 = 0, T_a = 0, T_b = 0;
    bool semi_implicit = false;
    double L_stab = 0;
    int jacobi_iters = 0;
};
struct DtWindow {
    double min_dx, S_scale, L_eff_total, D_total, dt_diff_total, L_eff_explicit, dt_diff,
        bulk_rate, dt_bulk, dt_max;
    bool ldg_elastic, anisotropic, semi_implicit_active;
};
DtWindow stable_dt(const DtInput& in);

// Correlation-length report and under-resolution test (QSR.cu:1615-1800).
// Prints the report to stdout; returns false when the guard is enabled and the
// coarsest-axis xi/dh falls below min_ratio.
struct Geometry;
bool correlation_length_guard(const Geometry& g, double a, double b, double c, double T,
                              double T_star, int modechoice, double kappa, double L1, double L2,
                              double L3, bool enabled, double min_ratio = 2.0);

// ---------------------------------------------------------------------------
// Droplet geometry (QSR.cu:34-185)
// ---------------------------------------------------------------------------
struct Geometry {
    int Nx = 1, Ny = 1, Nz = 1;
    double dx = 1, dy = 1, dz = 1;
    double cx = 0, cy = 0, cz = 0;  // centre (m)
    double rx = 1, ry = 1, rz = 1;  // semi-axes (m)
    double min_spacing = 1, min_radius = 1, shell_half = 1, core_radius = 1;
};
// radius_index <= 0 -> inscribed 0.5*(min N - 1); non-finite centre -> (N-1)/2.
Geometry make_geometry(int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double radius_index, const double center_index[3]);

struct Point {
    double x, y, z, r, sd;  // offsets from the centre, centre distance, signed distance
};
Point sample(const Geometry& g, int i, int j, int k);
void surface_normal(const Geometry& g, const Point& p, double n[3]);
void center_direction(const Point& p, double n[3]);
double surface_area(const Geometry& g);
double shell_thickness(const Geometry& g, uint64_t shell_count);

// The masks the step and the log pass need, as one k-interval per row (i, j)
// and set; empty intervals have lo > hi.  Built from the exact FP64 geometry so
// the device never re-derives a mask decision in lower precision.
//   E: evolve      sd <= band (band = shell_half if W > 0, else 0)   QSR.cu:108-111
//   O: near shell  sd <= shell_half           } shell = O \ I         QSR.cu:104-106
//   I: deep inside sd <  -shell_half          }
//   A: S average   sd <  -2*min_spacing                               QSR.cu:2289-2295
//   C: inside      sd <= 0                                             QSR.cu:100-102
//   K: core        r  <= 2*min_spacing   (radiality uses C \ K)        QSR.cu:1004
struct RowMasks {
    std::vector<int16_t> step;  // per row: e0 e1 o0 o1 n0 n1 0 0
    std::vector<int16_t> obs;   // per row: a0 a1 c0 c1 k0 k1 0 0
    uint64_t shell_count = 0;
    uint64_t active_count = 0;  // voxels that can change in a step
};
// Returns false (and fills err) if some row's set is not a single interval.
bool build_row_masks(const Geometry& g, double W, RowMasks& out, std::string& err);

// ---------------------------------------------------------------------------
// Seeded initial states (QSR.cu:1834-1904).  Q is AoS double, 5 per voxel.
// Uses std::mt19937 + std::uniform_real_distribution / std::normal_distribution
// in the reference's voxel and draw order; seed < 0 -> std::random_device.
// ---------------------------------------------------------------------------
void init_director(std::vector<double>& Q, const Geometry& g, double S0, int mode, long long seed);
void init_isotropic(std::vector<double>& Q, const Geometry& g, double amp, long long seed);

// ---------------------------------------------------------------------------
// Q-tensor text files, "i j k" + the full 3x3 tensor row-major (9 values), the
// format saveQTensorToFile writes (QSR.cu:1818-1832).  Reading follows the
// reference's own reader (QSRvis.py:2274-2304): '#' lines are comments, rows
// need >= 12 columns, rows outside the grid are ignored, off-diagonals are the
// mean of the two symmetric entries, Qzz is implied (traceless), voxels with no
// row stay 0.  Parsed in parallel over blocks of lines.
// ---------------------------------------------------------------------------
bool load_qtensor_text(const std::string& path, int Nx, int Ny, int Nz, std::vector<double>& Q, uint64_t& rows,
                       std::string& err);
// Writes Q (AoS, 5 per voxel) in i-major order with "%d %d %d" + 9 x " %.6e".
bool save_qtensor_text(const std::string& path, const double* Q, int Nx, int Ny, int Nz, std::string& err);

// ---------------------------------------------------------------------------
// key = value configuration (QSR.cu:1182-1297)
// ---------------------------------------------------------------------------
using KV = std::unordered_map<std::string, std::string>;
KV parse_kv_file(const std::string& path);
std::string kv_string(const KV& kv, const std::string& key, const std::string& def);
bool kv_bool(const KV& kv, const std::string& key, bool def);
long long kv_int(const KV& kv, const std::string& key, long long def);
double kv_double(const KV& kv, const std::string& key, double def);
std::string trim(const std::string& s);
std::string lower(std::string s);

}  // namespace qsrh
