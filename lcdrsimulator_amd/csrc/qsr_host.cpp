// qsr_host.cpp -- host physics, geometry, mask tables, seeded init, config.
#include "qsr_host.hpp"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <random>
#include <sstream>
#include <thread>

namespace qsrh {

static const double kNaN = std::numeric_limits<double>::quiet_NaN();

BulkABC bulk_abc(double a, double b, double c, double T, double T_star, int modechoice) {
    // A = 3 a (T - T*); std convention B = 9b/2, C = 9c/2; ravnik B = 27b/2, C = 9c.
    BulkABC k;
    k.A = 3.0 * a * (T - T_star);
    k.B = (modechoice == 1) ? 4.5 * b : 13.5 * b;
    k.C = (modechoice == 1) ? 4.5 * c : 9.0 * c;
    return k;
}

double bulk_f_uniaxial(const BulkABC& k, double S) {
    // trQ^2 = 2S^2/3, trQ^3 = 2S^3/9 for the uniaxial tensor.
    const double S2 = S * S;
    return (k.A / 3.0) * S2 - (2.0 * k.B / 27.0) * (S2 * S) + (k.C / 9.0) * (S2 * S2);
}

double s_eq(const BulkABC& k) {
    // Stationary points of f(S): 2C S^2 - B S + 3A = 0; keep the root with the
    // lowest f below f(0) = 0.
    if (!(k.C > 0.0)) return 0.0;
    const double disc = k.B * k.B - 24.0 * k.A * k.C;
    if (!(disc > 0.0)) return 0.0;
    const double root = std::sqrt(disc);
    double best_S = 0.0, best_f = 0.0;
    for (double S : {(k.B - root) / (4.0 * k.C), (k.B + root) / (4.0 * k.C)}) {
        if (!(S > 0.0)) continue;
        const double f = bulk_f_uniaxial(k, S);
        if (f < best_f) {
            best_f = f;
            best_S = S;
        }
    }
    return best_S;
}

double a_coexistence(double B, double C) { return (C > 0.0) ? B * B / (27.0 * C) : kNaN; }
double a_spinodal_nematic(double B, double C) { return (C > 0.0) ? B * B / (24.0 * C) : kNaN; }
double t_from_a(double A, double a, double T_star) { return (a > 0.0) ? T_star + A / (3.0 * a) : kNaN; }

double shell_order(bool custom, double custom_S, double a, double b, double c, double T,
                   double T_star, int modechoice) {
    if (custom) return custom_S;
    const double S = s_eq(bulk_abc(a, b, c, T, T_star, modechoice));
    return S > 0.0 ? S : 0.0;
}

DtWindow stable_dt(const DtInput& in) {
    DtWindow w{};
    w.min_dx = std::min({in.dx, in.dy, in.dz});
    w.S_scale = std::isfinite(in.S_ref_used) ? std::abs(in.S_ref_used) : std::max(std::abs(in.S0), 0.5);
    for (double T : {in.T_a, in.T_b}) {
        const double S = std::abs(shell_order(false, 0.0, in.a, in.b, in.c, T, in.T_star, in.modechoice));
        if (std::isfinite(S) && S > 0.0) w.S_scale = std::max(w.S_scale, S);
    }
    w.S_scale = std::max(w.S_scale, 1.0);
    w.ldg_elastic = (in.L1 != 0.0 || in.L2 != 0.0 || in.L3 != 0.0);
    w.anisotropic = (in.L2 != 0.0 || in.L3 != 0.0);
    auto ldg_envelope = [&](double l1_part) {
        return std::max({l1_part, std::abs(in.L2), std::abs(in.L1 + in.L2),
                         std::abs(in.L1 + 2.0 * in.L2), std::abs(in.L3) * w.S_scale});
    };
    w.L_eff_total = w.ldg_elastic ? ldg_envelope(std::abs(in.L1)) : std::abs(in.kappa);
    if (!(w.L_eff_total > 0.0)) w.L_eff_total = 1e-12;
    w.D_total = w.L_eff_total / in.gamma;
    if (!(w.D_total > 0.0) || !std::isfinite(w.D_total)) w.D_total = 1e-12;
    const double aniso_factor = w.anisotropic ? 0.25 : 1.0;
    w.dt_diff_total = 0.5 * (w.min_dx * w.min_dx) / (6.0 * w.D_total) * aniso_factor;
    w.L_eff_explicit = w.L_eff_total;
    w.dt_diff = w.dt_diff_total;
    if (in.semi_implicit && in.L_stab > 0.0 && in.jacobi_iters > 0) {
        const double L_base = w.ldg_elastic ? in.L1 : in.kappa;
        const double L_rem = std::abs(L_base - ((L_base >= 0.0) ? in.L_stab : -in.L_stab));
        w.L_eff_explicit = w.ldg_elastic ? ldg_envelope(L_rem) : L_rem;
        w.dt_diff = (w.L_eff_explicit > 0.0)
                        ? 0.5 * (w.min_dx * w.min_dx) / (6.0 * (w.L_eff_explicit / in.gamma)) * aniso_factor
                        : std::numeric_limits<double>::infinity();
        w.semi_implicit_active = true;
    }
    w.bulk_rate = 0.0;
    for (double T : {in.T_a, in.T_b}) {
        const BulkABC k = bulk_abc(in.a, in.b, in.c, T, in.T_star, in.modechoice);
        w.bulk_rate = std::max(w.bulk_rate, std::abs(k.A) + std::abs(k.B) * w.S_scale +
                                                std::abs(k.C) * w.S_scale * w.S_scale);
    }
    w.dt_bulk = (w.bulk_rate > 0.0) ? 0.02 * in.gamma / w.bulk_rate : std::numeric_limits<double>::infinity();
    w.dt_max = std::min(w.dt_diff, w.dt_bulk);
    if (!(w.dt_max > 0.0) || !std::isfinite(w.dt_max)) w.dt_max = w.dt_diff;
    return w;
}

bool correlation_length_guard(const Geometry& g, double a, double b, double c, double T,
                              double T_star, int modechoice, double kappa, double L1, double L2,
                              double L3, bool enabled, double min_ratio) {
    const BulkABC k = bulk_abc(a, b, c, T, T_star, modechoice);
    const double absA_lin = std::abs(k.A);
    const double S = s_eq(k);
    // f''(S) of the uniaxial reduction at S_eq
    const double fpp = (S > 0.0) ? (2.0 * k.A / 3.0 - (4.0 * k.B / 9.0) * S + (4.0 * k.C / 3.0) * S * S) : kNaN;
    const double absA_nem = std::isfinite(fpp) ? std::abs(fpp) : kNaN;
    const bool ldg = (L1 != 0.0 || L2 != 0.0 || L3 != 0.0);
    double base_min, base_max, nem_min, nem_max, l3_env = 0.0;
    bool fallback = false;
    if (!ldg) {
        base_min = base_max = nem_min = nem_max = std::abs(kappa);
    } else {
        std::vector<double> pos;
        for (double v : {L1, L1 + L2, L1 + 2.0 * L2})
            if (std::isfinite(v) && v > 0.0) pos.push_back(v);
        if (pos.empty()) {
            base_min = base_max = nem_min = nem_max =
                std::max({std::abs(L1), std::abs(L1 + L2), std::abs(L1 + 2.0 * L2), 1e-12});
            fallback = true;
        } else {
            base_min = *std::min_element(pos.begin(), pos.end());
            base_max = *std::max_element(pos.begin(), pos.end());
            l3_env = (S > 0.0 && std::isfinite(S)) ? std::abs(L3) * S : 0.0;
            nem_min = base_min;
            nem_max = base_max;
            if (l3_env > 0.0) {
                nem_min = std::max(base_min - l3_env, 1e-12);
                nem_max = base_max + l3_env;
            }
        }
    }
    auto xi_of = [](double L, double absA) { return (L <= 0.0 || absA <= 0.0) ? kNaN : std::sqrt(L / absA); };
    const double xi_lin_min = xi_of(base_min, absA_lin), xi_lin_max = xi_of(base_max, absA_lin);
    const double xi_nem_min = fallback ? xi_of(base_min, absA_nem) : xi_of(nem_min, absA_nem);
    const double xi_nem_max = fallback ? xi_of(base_max, absA_nem) : xi_of(nem_max, absA_nem);
    const double xi_guard = std::isfinite(xi_nem_min) ? xi_nem_min : xi_lin_min;
    const double xi_conf = std::isfinite(xi_nem_max) ? xi_nem_max : (std::isfinite(xi_lin_max) ? xi_lin_max : xi_guard);
    auto ratio = [&](double d) { return (std::isfinite(xi_guard) && d > 0.0) ? xi_guard / d : 0.0; };
    const double worst = std::min({ratio(g.dx), ratio(g.dy), ratio(g.dz)});
    std::cout << "\n--- Correlation Length (xi) Estimate ---\n";
    if (ldg) {
        std::cout << "Elastic stiffness band for xi (LdG): base=[" << base_min << ", " << base_max << "] J/m";
        if (l3_env > 0.0)
            std::cout << ", nematic=[" << nem_min << ", " << nem_max << "] J/m using +- |L3| S_eq = " << l3_env << " J/m";
        std::cout << "\n";
    } else {
        std::cout << "Elastic stiffness for xi: " << base_min << " J/m (kappa)\n";
    }
    std::cout << "Bulk convention = " << (modechoice == 1 ? "std" : "ravnik") << "\n";
    std::cout << "Bulk stiffness (linear) |A_lin| = |A| = |3 a (T-T*)| = " << absA_lin << " J/m^3\n";
    if (std::isfinite(xi_lin_min)) std::cout << "xi_lin band ~ [" << xi_lin_min << ", " << xi_lin_max << "] m\n";
    if (std::isfinite(xi_nem_min)) std::cout << "xi_nem band ~ [" << xi_nem_min << ", " << xi_nem_max << "] m\n";
    if (std::isfinite(xi_guard)) {
        std::cout << "xi_guard = " << xi_guard << " m (worst axis ratio = " << worst << ")\n";
        std::cout << "Geometry confinement: min R/xi_conf = "
                  << std::min({g.rx, g.ry, g.rz}) / xi_conf << "\n";
    } else {
        std::cout << "xi estimate unavailable (check elastic stiffness and that T is not exactly T*)\n";
    }
    if (!enabled) return true;
    if (std::isfinite(xi_guard) && worst < min_ratio) {
        std::cout << "\n[WARN] xi is under-resolved on the coarsest grid axis (worst xi/dh ~ " << worst
                  << " < " << min_ratio << ").\n       Non-interactive run: aborting instead of prompting.\n";
        return false;
    }
    return true;
}

Geometry make_geometry(int Nx, int Ny, int Nz, double dx, double dy, double dz,
                       double radius_index, const double center_index[3]) {
    Geometry g;
    g.Nx = Nx; g.Ny = Ny; g.Nz = Nz;
    g.dx = dx; g.dy = dy; g.dz = dz;
    const int N[3] = {Nx, Ny, Nz};
    const double d[3] = {dx, dy, dz};
    double* cen[3] = {&g.cx, &g.cy, &g.cz};
    for (int a = 0; a < 3; ++a) {
        const double ci = (center_index && std::isfinite(center_index[a])) ? center_index[a]
                                                                            : 0.5 * (double(N[a]) - 1.0);
        *cen[a] = ci * d[a];
    }
    const double ri = (radius_index > 0.0) ? radius_index : 0.5 * (double(std::min({Nx, Ny, Nz})) - 1.0);
    g.rx = std::max(ri * dx, 0.5 * dx);
    g.ry = std::max(ri * dy, 0.5 * dy);
    g.rz = std::max(ri * dz, 0.5 * dz);
    g.min_spacing = std::min({dx, dy, dz});
    g.min_radius = std::min({g.rx, g.ry, g.rz});
    g.shell_half = g.min_spacing;
    g.core_radius = g.min_spacing;
    return g;
}

Point sample(const Geometry& g, int i, int j, int k) {
    // Signed distance to the ellipsoid surface, first-order: (rho - 1) / |grad rho|.
    Point p;
    p.x = i * g.dx - g.cx;
    p.y = j * g.dy - g.cy;
    p.z = k * g.dz - g.cz;
    p.r = std::sqrt(p.x * p.x + p.y * p.y + p.z * p.z);
    const double ux = p.x / g.rx, uy = p.y / g.ry, uz = p.z / g.rz;
    const double rho2 = ux * ux + uy * uy + uz * uz;
    if (rho2 <= 1e-30) {
        p.sd = -g.min_radius;
        return p;
    }
    const double rho = std::sqrt(rho2);
    const double gx = p.x / (g.rx * g.rx), gy = p.y / (g.ry * g.ry), gz = p.z / (g.rz * g.rz);
    const double gn = std::sqrt(gx * gx + gy * gy + gz * gz) / rho;
    p.sd = (gn > 1e-30) ? (rho - 1.0) / gn : p.r - g.min_radius;
    return p;
}

void center_direction(const Point& p, double n[3]) {
    if (p.r <= 1e-30) {
        n[0] = 1.0; n[1] = 0.0; n[2] = 0.0;
        return;
    }
    n[0] = p.x / p.r; n[1] = p.y / p.r; n[2] = p.z / p.r;
}

void surface_normal(const Geometry& g, const Point& p, double n[3]) {
    const double gx = p.x / (g.rx * g.rx), gy = p.y / (g.ry * g.ry), gz = p.z / (g.rz * g.rz);
    const double gn = std::sqrt(gx * gx + gy * gy + gz * gz);
    if (gn <= 1e-30) {
        center_direction(p, n);
        return;
    }
    n[0] = gx / gn; n[1] = gy / gn; n[2] = gz / gn;
}

double surface_area(const Geometry& g) {
    // Thomsen's approximation, p = 1.6075.
    const double p = 1.6075;
    const double a = std::pow(g.rx, p), b = std::pow(g.ry, p), c = std::pow(g.rz, p);
    const double m = (a * b + a * c + b * c) / 3.0;
    if (!(m > 0.0)) return kNaN;
    return 4.0 * kPi * std::pow(m, 1.0 / p);
}

double shell_thickness(const Geometry& g, uint64_t shell_count) {
    const double nominal = 2.0 * g.shell_half;
    if (shell_count == 0) return nominal;
    const double area = surface_area(g);
    if (!(area > 0.0)) return nominal;
    const double t = double(shell_count) * g.dx * g.dy * g.dz / area;
    return (std::isfinite(t) && t > 0.0) ? t : nominal;
}

namespace {
// Collapse a row's membership flags into [lo, hi]; false if not contiguous.
bool to_interval(const std::vector<uint8_t>& in, int16_t& lo, int16_t& hi) {
    int first = -1, last = -2;
    for (int k = 0; k < int(in.size()); ++k)
        if (in[k]) {
            if (first < 0) first = k;
            else if (last != k - 1) return false;
            last = k;
        }
    if (first < 0) {
        lo = 1;
        hi = 0;
    } else {
        lo = int16_t(first);
        hi = int16_t(last);
    }
    return true;
}
}  // namespace

bool build_row_masks(const Geometry& g, double W, RowMasks& out, std::string& err) {
    if (g.Nz > 32767) {
        err = "Nz too large for 16-bit row intervals";
        return false;
    }
    const size_t rows = size_t(g.Nx) * g.Ny;
    out.step.assign(rows * 8, 0);
    out.obs.assign(rows * 8, 0);
    out.shell_count = 0;
    out.active_count = 0;
    const double band = (W > 0.0) ? g.shell_half : 0.0;
    const double excl = 2.0 * g.min_spacing;
    std::vector<uint8_t> sE(g.Nz), sO(g.Nz), sI(g.Nz), sA(g.Nz), sC(g.Nz), sK(g.Nz);
    for (int i = 0; i < g.Nx; ++i)
        for (int j = 0; j < g.Ny; ++j) {
            for (int k = 0; k < g.Nz; ++k) {
                const Point p = sample(g, i, j, k);
                sE[k] = p.sd <= band;
                sO[k] = p.sd <= g.shell_half;
                sI[k] = p.sd < -g.shell_half;
                sA[k] = p.sd < -excl;
                sC[k] = p.sd <= 0.0;
                sK[k] = p.r <= excl;
                const bool shell = std::abs(p.sd) <= g.shell_half;
                if (shell != (sO[k] && !sI[k])) {
                    err = "shell set is not O minus I";
                    return false;
                }
                out.shell_count += shell;
                // a voxel changes if it evolves (and is not a frozen strong shell)
                // or is re-imposed by the strong boundary condition
                const bool evolves = sE[k] && !(W == 0.0 && shell);
                const bool bc = shell && !(W > 0.0);
                out.active_count += (evolves || bc);
            }
            int16_t* s = &out.step[(size_t(i) * g.Ny + j) * 8];
            int16_t* o = &out.obs[(size_t(i) * g.Ny + j) * 8];
            if (!to_interval(sE, s[0], s[1]) || !to_interval(sO, s[2], s[3]) || !to_interval(sI, s[4], s[5]) ||
                !to_interval(sA, o[0], o[1]) || !to_interval(sC, o[2], o[3]) || !to_interval(sK, o[4], o[5])) {
                err = "droplet mask row (" + std::to_string(i) + "," + std::to_string(j) +
                      ") is not a single k-interval";
                return false;
            }
        }
    return true;
}

static std::mt19937 seeded_engine(long long seed) {
    if (seed >= 0) return std::mt19937(static_cast<std::mt19937::result_type>(seed));
    return std::mt19937(std::random_device{}());
}

void init_director(std::vector<double>& Q, const Geometry& g, double S0, int mode, long long seed) {
    const size_t n = size_t(g.Nx) * g.Ny * g.Nz;
    Q.assign(5 * n, 0.0);
    std::mt19937 eng = seeded_engine(seed);
    std::uniform_real_distribution<double> theta_dist(0.0, kPi), phi_dist(0.0, 2.0 * kPi);
    size_t id = 0;
    for (int i = 0; i < g.Nx; ++i)
        for (int j = 0; j < g.Ny; ++j)
            for (int k = 0; k < g.Nz; ++k, ++id) {
                const Point p = sample(g, i, j, k);
                if (!(p.sd <= 0.0)) continue;
                double S = S0;
                if (p.r < g.core_radius) S *= p.r / g.core_radius;
                double n3[3];
                if (mode == 0) {
                    const double th = theta_dist(eng);
                    const double ph = phi_dist(eng);
                    n3[0] = std::sin(th) * std::cos(ph);
                    n3[1] = std::sin(th) * std::sin(ph);
                    n3[2] = std::cos(th);
                } else {
                    center_direction(p, n3);
                }
                double* q = &Q[5 * id];
                q[0] = S * (n3[0] * n3[0] - 1.0 / 3.0);
                q[1] = S * (n3[0] * n3[1]);
                q[2] = S * (n3[0] * n3[2]);
                q[3] = S * (n3[1] * n3[1] - 1.0 / 3.0);
                q[4] = S * (n3[1] * n3[2]);
            }
}

void init_isotropic(std::vector<double>& Q, const Geometry& g, double amp, long long seed) {
    const size_t n = size_t(g.Nx) * g.Ny * g.Nz;
    Q.assign(5 * n, 0.0);
    if (amp <= 0.0) return;
    std::mt19937 eng = seeded_engine(seed);
    std::normal_distribution<double> dist(0.0, amp);
    size_t id = 0;
    for (int i = 0; i < g.Nx; ++i)
        for (int j = 0; j < g.Ny; ++j)
            for (int k = 0; k < g.Nz; ++k, ++id) {
                if (!(sample(g, i, j, k).sd <= 0.0)) continue;
                for (int c = 0; c < 5; ++c) Q[5 * id + c] = dist(eng);
            }
}

namespace {

// Parses [b, e): whole lines.  Returns false on a malformed data row.
bool parse_qtensor_lines(const char* b, const char* e, int Nx, int Ny, int Nz, double* Q, uint64_t& rows,
                         std::string& err) {
    const char* p = b;
    while (p < e) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', size_t(e - p)));
        const char* le = nl ? nl : e;
        const char* q = p;
        while (q < le && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
        if (q < le && *q != '#') {
            double v[12];
            int nv = 0;
            while (q < le && nv < 12) {
                while (q < le && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
                if (q >= le) break;
                auto r = std::from_chars(q, le, v[nv]);
                if (r.ec != std::errc()) {
                    err = "malformed number in row: " + std::string(p, size_t(std::min<ptrdiff_t>(le - p, 120)));
                    return false;
                }
                q = r.ptr;
                ++nv;
            }
            if (nv < 12) {
                err = "row with fewer than 12 columns: " + std::string(p, size_t(std::min<ptrdiff_t>(le - p, 120)));
                return false;
            }
            const long i = long(v[0]), j = long(v[1]), k = long(v[2]);
            if (i >= 0 && i < Nx && j >= 0 && j < Ny && k >= 0 && k < Nz) {
                double* o = Q + 5 * (size_t(k) + size_t(Nz) * (size_t(j) + size_t(Ny) * size_t(i)));
                o[0] = v[3];
                o[1] = 0.5 * (v[4] + v[6]);
                o[2] = 0.5 * (v[5] + v[9]);
                o[3] = v[7];
                o[4] = 0.5 * (v[8] + v[10]);
            }
            ++rows;
        }
        p = nl ? nl + 1 : e;
    }
    return true;
}

}  // namespace

bool load_qtensor_text(const std::string& path, int Nx, int Ny, int Nz, std::vector<double>& Q, uint64_t& rows,
                       std::string& err) {
    rows = 0;
    if (Nx < 1 || Ny < 1 || Nz < 1) {
        err = "bad grid size";
        return false;
    }
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    Q.assign(size_t(5) * size_t(Nx) * size_t(Ny) * size_t(Nz), 0.0);
    const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    constexpr size_t kBlock = size_t(64) << 20;
    std::vector<char> buf;
    std::string carry;
    bool ok = true;
    for (;;) {
        buf.assign(carry.begin(), carry.end());
        const size_t have = buf.size();
        buf.resize(have + kBlock);
        const size_t got = std::fread(buf.data() + have, 1, kBlock, f);
        buf.resize(have + got);
        const bool last = got < kBlock;
        size_t end = buf.size();
        carry.clear();
        if (!last) {
            const void* nl = nullptr;
            for (size_t x = buf.size(); x > 0 && !nl; --x)
                if (buf[x - 1] == '\n') nl = &buf[x - 1];
            if (!nl) {  // a single line longer than a block: keep reading
                carry.assign(buf.begin(), buf.end());
                continue;
            }
            end = size_t(static_cast<const char*>(nl) - buf.data()) + 1;
            carry.assign(buf.begin() + long(end), buf.end());
        }
        // split [0, end) at line boundaries into nth parts
        std::vector<size_t> cut{0};
        for (unsigned t = 1; t < nth; ++t) {
            size_t x = end * t / nth;
            while (x > 0 && x < end && buf[x - 1] != '\n') ++x;
            cut.push_back(std::max(cut.back(), x));
        }
        cut.push_back(end);
        std::vector<uint64_t> nrow(nth, 0);
        std::vector<std::string> errs(nth);
        std::vector<char> good(nth, 1);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; ++t)
            th.emplace_back([&, t] {
                good[t] = parse_qtensor_lines(buf.data() + cut[t], buf.data() + cut[t + 1], Nx, Ny, Nz, Q.data(),
                                              nrow[t], errs[t]);
            });
        for (auto& x : th) x.join();
        for (unsigned t = 0; t < nth; ++t) {
            rows += nrow[t];
            if (!good[t] && ok) {
                ok = false;
                err = errs[t];
            }
        }
        if (!ok || last) break;
    }
    std::fclose(f);
    if (ok && !carry.empty()) {
        uint64_t n = 0;
        ok = parse_qtensor_lines(carry.data(), carry.data() + carry.size(), Nx, Ny, Nz, Q.data(), n, err);
        rows += n;
    }
    return ok;
}

bool save_qtensor_text(const std::string& path, const double* Q, int Nx, int Ny, int Nz, std::string& err) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) {
        err = "cannot write " + path;
        return false;
    }
    const unsigned nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    for (int i0 = 0; i0 < Nx; i0 += int(nth)) {
        const int n = std::min(int(nth), Nx - i0);
        std::vector<std::string> out(static_cast<size_t>(n));
        std::vector<std::thread> th;
        for (int t = 0; t < n; ++t)
            th.emplace_back([&, t] {
                const int i = i0 + t;
                char line[512];
                std::string& o = out[size_t(t)];
                o.reserve(size_t(Ny) * size_t(Nz) * 140);
                for (int j = 0; j < Ny; ++j)
                    for (int k = 0; k < Nz; ++k) {
                        const double* q = Q + 5 * (size_t(k) + size_t(Nz) * (size_t(j) + size_t(Ny) * size_t(i)));
                        const double zz = -q[0] - q[3];
                        const int m = std::snprintf(line, sizeof line,
                                                    "%d %d %d %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e\n", i, j,
                                                    k, q[0], q[1], q[2], q[1], q[3], q[4], q[2], q[4], zz);
                        o.append(line, size_t(m));
                    }
            });
        for (auto& x : th) x.join();
        for (auto& o : out) std::fwrite(o.data(), 1, o.size(), f);
    }
    const bool ok = std::fclose(f) == 0;
    if (!ok) err = "write failed: " + path;
    return ok;
}

std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace(static_cast<unsigned char>(s[b]))) ++b;
    while (e > b && std::isspace(static_cast<unsigned char>(s[e - 1]))) --e;
    return s.substr(b, e - b);
}

std::string lower(std::string s) {
    for (char& ch : s) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    return s;
}

KV parse_kv_file(const std::string& path) {
    // "key = value" lines, '#' starts a comment, last duplicate wins.
    KV kv;
    if (path.empty()) return kv;
    std::ifstream f(path);
    if (!f) {
        std::cerr << "[Config] Could not open config file: " << path << std::endl;
        return kv;
    }
    std::string line;
    while (std::getline(f, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        line = trim(line);
        const size_t eq = line.find('=');
        if (line.empty() || eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        if (!key.empty()) kv[key] = trim(line.substr(eq + 1));
    }
    return kv;
}

std::string kv_string(const KV& kv, const std::string& key, const std::string& def) {
    auto it = kv.find(key);
    return it == kv.end() ? def : it->second;
}

bool kv_bool(const KV& kv, const std::string& key, bool def) {
    auto it = kv.find(key);
    if (it == kv.end() || it->second.empty()) return def;
    const std::string v = lower(trim(it->second));
    if (v == "1" || v == "true" || v == "yes" || v == "y" || v == "on") return true;
    if (v == "0" || v == "false" || v == "no" || v == "n" || v == "off") return false;
    return def;
}

long long kv_int(const KV& kv, const std::string& key, long long def) {
    // Integers go through long double so "1e6" is accepted; out-of-range or
    // trailing garbage keeps the default.
    auto it = kv.find(key);
    if (it == kv.end() || it->second.empty()) return def;
    try {
        size_t used = 0;
        const long double v = std::stold(it->second, &used);
        while (used < it->second.size() && std::isspace(static_cast<unsigned char>(it->second[used]))) ++used;
        if (used != it->second.size()) return def;
        const long double r = std::llround(v);
        if (r < (long double)std::numeric_limits<long long>::min() ||
            r > (long double)std::numeric_limits<long long>::max())
            return def;
        return static_cast<long long>(r);
    } catch (...) {
        return def;
    }
}

double kv_double(const KV& kv, const std::string& key, double def) {
    auto it = kv.find(key);
    if (it == kv.end() || it->second.empty()) return def;
    std::istringstream iss(it->second);
    double v;
    if (!(iss >> v)) return def;
    return v;
}

}  // namespace qsrh
