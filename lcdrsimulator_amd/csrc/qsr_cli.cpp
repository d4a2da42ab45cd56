// qsr_cli.cpp -- `QSR_hip --config <file>`: the command-line driver.
//
// Same contract as the reference binary (QSR.cu:1910-3503): key = value config
// (same keys and defaults), sim_mode 1 (single temperature relaxation),
// 2 (temperature sweep) and 3 (quench), the same output files and headers
// (free_energy_vs_iteration.dat / energy_components_vs_iteration.dat,
// output_temp_sweep/summary.dat, <out_dir>/quench_log.dat, nematic_field_*.dat,
// Qtensor_output_*.dat), the machine-parsed stdout lines and exit codes
// (0 ok, 1 guard abort).  The step runs on the GPU through the C ABI.
//
// Additions (defaults keep the reference behaviour): droplet_radius and
// droplet_center_{x,y,z} (voxels), save_iteration_snapshots (sim_mode 1 writes
// output/nematic_field_iter_*.dat at every log point when true),
// early_stop_min_alignment_time (the 0.5*tau_align gate of QSR.cu:3193),
// noise_sigma / noise_seed (Philox noise per step), step_graph_chunk,
// init_qtensor_file (warm start / resume from a Qtensor_output_*.dat file).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/qsr_amd.h"
#include "qsr_host.hpp"

namespace fs = std::filesystem;
using qsrh::KV;

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

struct Cfg {
    KV kv;
    bool interactive = false;
    std::string get_s(const std::string& key, const std::string& def, const char* prompt = nullptr) const {
        std::string v = qsrh::kv_string(kv, key, def);
        if (interactive) v = ask(prompt ? prompt : key.c_str(), v);
        return v;
    }
    double get_d(const std::string& key, double def, const char* prompt = nullptr) const {
        double v = qsrh::kv_double(kv, key, def);
        if (interactive) {
            const std::string a = ask(prompt ? prompt : key.c_str(), num(v));
            KV tmp{{key, a}};
            v = qsrh::kv_double(tmp, key, v);
        }
        return v;
    }
    long long get_ll(const std::string& key, long long def, const char* prompt = nullptr) const {
        long long v = qsrh::kv_int(kv, key, def);
        if (interactive) {
            KV tmp{{key, ask(prompt ? prompt : key.c_str(), std::to_string(v))}};
            v = qsrh::kv_int(tmp, key, v);
        }
        return v;
    }
    int get_i(const std::string& key, int def, const char* prompt = nullptr) const {
        const long long v = get_ll(key, def, prompt);
        if (v < std::numeric_limits<int>::min() || v > std::numeric_limits<int>::max()) return def;
        return int(v);
    }
    bool get_b(const std::string& key, bool def, const char* prompt = nullptr) const {
        bool v = qsrh::kv_bool(kv, key, def);
        if (interactive) {
            const std::string a = ask(prompt ? prompt : key.c_str(), v ? "y" : "n");
            if (!a.empty() && (a[0] == 'y' || a[0] == 'Y')) v = true;
            else if (!a.empty() && (a[0] == 'n' || a[0] == 'N')) v = false;
        }
        return v;
    }
    static std::string num(double v) {
        std::ostringstream o;
        o << v;
        return o.str();
    }
    static std::string ask(const char* prompt, const std::string& def) {
        std::cout << prompt << " [" << def << "]: " << std::flush;
        std::string line;
        if (!std::getline(std::cin, line) || line.empty()) return def;
        return line;
    }
};

int check(int rc, const char* what) {
    if (rc != QSR_OK) {
        std::cerr << "[QSR_hip] " << what << " failed (" << rc << "): " << qsr_last_error() << std::endl;
    }
    return rc;
}

// Text writers with the reference formats (QSR.cu:1802-1832), formatted in
// parallel by i-plane and written in order.
template <class Fmt>
void write_planes(const std::string& path, int Nx, const char* header, Fmt&& fmt_plane) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    if (header) std::fputs(header, f);
    const int nth = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    for (int i0 = 0; i0 < Nx; i0 += nth) {
        const int n = std::min(nth, Nx - i0);
        std::vector<std::string> buf(n);
        std::vector<std::thread> th;
        for (int t = 0; t < n; ++t) th.emplace_back([&, t] { fmt_plane(i0 + t, buf[t]); });
        for (auto& x : th) x.join();
        for (auto& s : buf) std::fwrite(s.data(), 1, s.size(), f);
    }
    std::fclose(f);
}

void save_q(const std::vector<double>& Q, int Nx, int Ny, int Nz, const std::string& path) {
    std::string err;
    if (!qsrh::save_qtensor_text(path, Q.data(), Nx, Ny, Nz, err)) std::cerr << "[Output] " << err << std::endl;
}

void save_nematic(const std::vector<double>& nf, int Nx, int Ny, int Nz, const std::string& path) {
    write_planes(path, Nx, "# i j k S nx ny nz\n", [&](int i, std::string& out) {
        char line[256];
        out.reserve(size_t(Ny) * Nz * 64);
        for (int j = 0; j < Ny; ++j)
            for (int k = 0; k < Nz; ++k) {
                const double* v = &nf[4 * (size_t(k) + size_t(Nz) * (size_t(j) + size_t(Ny) * i))];
                const int n = std::snprintf(line, sizeof line, "%d %d %d %.6e %.6e %.6e %.6e\n", i, j, k, v[0], v[1],
                                            v[2], v[3]);
                out.append(line, size_t(n));
            }
    });
}

struct Run {
    void* ctx = nullptr;
    qsr_params p{};
    size_t n = 0;
    ~Run() {
        if (ctx) qsr_destroy(ctx);
    }
    int save_final(const std::string& nem_path, const std::string& q_path) {
        std::vector<double> Q(5 * n), nf(4 * n);
        if (check(qsr_download_q(ctx, 0, Q.data(), n), "download")) return 1;
        if (check(qsr_download_nematic(ctx, 0, nf.data(), n), "nematic")) return 1;
        save_nematic(nf, p.Nx, p.Ny, p.Nz, nem_path);
        save_q(Q, p.Nx, p.Ny, p.Nz, q_path);
        return 0;
    }
    int save_nematic_only(const std::string& path) {
        std::vector<double> nf(4 * n);
        if (check(qsr_download_nematic(ctx, 0, nf.data(), n), "nematic")) return 1;
        save_nematic(nf, p.Nx, p.Ny, p.Nz, path);
        return 0;
    }
    int save_q_only(const std::string& path) {
        std::vector<double> Q(5 * n);
        if (check(qsr_download_q(ctx, 0, Q.data(), n), "download")) return 1;
        save_q(Q, p.Nx, p.Ny, p.Nz, path);
        return 0;
    }
};

void print_help() {
    std::cout << "QSR_hip options:\n"
              << "  --config <path>       Run non-interactively using key=value config file\n"
              << "  --interactive         When used with --config, still prompt (config values become defaults)\n"
              << "  --help                Show this help\n";
}

void print_dt(const qsrh::DtWindow& w) {
    std::cout << "\nMaximum stable dt estimates:\n"
              << "  dt_diff (elastic) ~= " << w.dt_diff << " s\n"
              << "  dt_diff_total     ~= " << w.dt_diff_total << " s"
              << (w.semi_implicit_active ? " (before semi-implicit)" : "") << "\n"
              << "  dt_bulk (bulk)    ~= " << w.dt_bulk << " s\n"
              << "  dt_max            = " << w.dt_max << " s" << std::endl;
}

}  // namespace

int main(int argc, char** argv) {
    std::string config_path;
    bool interactive = true;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i] ? argv[i] : "";
        if (a == "--help" || a == "-h") {
            print_help();
            return 0;
        }
        if (a == "--config" && i + 1 < argc) {
            config_path = argv[++i];
            interactive = false;
            continue;
        }
        if (a == "--interactive") interactive = true;
    }
    Cfg cfg;
    cfg.kv = qsrh::parse_kv_file(config_path);
    cfg.interactive = interactive;
    std::cout << "Q-tensor evolution simulation (HIP, MI355X)\n" << std::endl;
    if (!interactive)
        std::cout << "[Config] Running non-interactively" << (config_path.empty() ? " (no config path?)" : " with " + config_path)
                  << "\n";

    // ---- parameters in the reference's order (QSR.cu:1958-2139) --------------------------------
    int mode = std::clamp(cfg.get_i("init_mode", 1, "Select initialization mode: [0] random, [1] radial, [2] isotropic+noise"), 0, 2);
    const long long random_seed = cfg.get_ll("random_seed", -1, "Random seed for stochastic initialization (-1 = random_device)");
    qsr_params p{};
    p.Nx = cfg.get_i("Nx", 100, "Enter grid size Nx");
    p.Ny = cfg.get_i("Ny", 100, "Enter grid size Ny");
    p.Nz = cfg.get_i("Nz", 100, "Enter grid size Nz");
    p.dx = cfg.get_d("dx", 1e-8, "Enter dx (m)");
    p.dy = cfg.get_d("dy", 1e-8, "Enter dy (m)");
    p.dz = cfg.get_d("dz", 1e-8, "Enter dz (m)");
    p.a = cfg.get_d("a", 0.044e6, "Enter a (J/m^3/K)");
    p.b = cfg.get_d("b", 1.413e6, "Enter b (J/m^3)");
    p.c = cfg.get_d("c", 1.153e6, "Enter c (J/m^3)");
    const double T = cfg.get_d("T", 300.0, "Enter T (K)");
    p.T_star = cfg.get_d("T_star", 308.0, "Enter T* (K) [A=0, isotropic spinodal]");
    p.modechoice = cfg.get_i("bulk_modechoice", 1, "Enter Bulk Energy convention (1=std, 2=ravnik)");
    const qsrh::BulkABC abc = qsrh::bulk_abc(p.a, p.b, p.c, T, p.T_star, p.modechoice);
    const double S_eq = qsrh::s_eq(abc);
    {
        const qsrh::BulkABC k0 = qsrh::bulk_abc(p.a, p.b, p.c, T, p.T_star, p.modechoice);
        const double T_NI = qsrh::t_from_a(qsrh::a_coexistence(k0.B, k0.C), p.a, p.T_star);
        const double T_sp = qsrh::t_from_a(qsrh::a_spinodal_nematic(k0.B, k0.C), p.a, p.T_star);
        std::cout << "\nBulk convention: " << (p.modechoice == 1 ? "std" : "ravnik") << "\nComputed S_c (heuristic) = "
                  << p.b / (2.0 * p.c) << ", S_eq(T=" << T << "K) = " << S_eq << std::endl;
        if (std::isfinite(T_NI) && std::isfinite(T_sp))
            std::cout << "Bulk transition estimates (uniaxial, homogeneous):\n  T* (A=0, isotropic spinodal) = " << p.T_star
                      << " K\n  T_NI (coexistence, global min) ~ " << T_NI << " K\n  T_spin,N (nematic disappears)  ~ "
                      << T_sp << " K\n";
    }
    const double S0 = cfg.get_d("S0", S_eq > 0 ? S_eq : 0.5, "Enter initial S0 (initial condition only)");
    p.kappa = cfg.get_d("kappa", 6.5e-12, "Enter kappa (J/m)");
    double S_ref_used = kNaN;
    if (cfg.get_b("use_frank_map", false, "Use Frank-to-LdG mapping with K1, K2, K3?")) {
        const double K1 = cfg.get_d("K1", 6.5e-12), K2 = cfg.get_d("K2", 4.0e-12), K3 = cfg.get_d("K3", 8e-12);
        double S_ref = cfg.get_d("S_ref", S_eq > 0.0 ? S_eq : (S0 > 0.0 ? S0 : p.b / (2.0 * p.c)));
        if (S_ref <= 1e-12) S_ref = 0.5;
        S_ref_used = S_ref;
        // L2 = (K1-K2)/S^2, L3 = (9/4)(K3-K1)/S^3, L1 = (K1+K2-K3)/(2 S^2)  (QSR.cu:2026-2040)
        const double S2 = S_ref * S_ref, S3 = S2 * S_ref;
        p.L2 = (K1 - K2) / S2;
        p.L3 = 2.25 * (K3 - K1) / S3;
        p.L1 = (K1 + K2 - K3) / (2.0 * S2);
        p.kappa = 0.0;
        std::cout << "Mapped L1=" << p.L1 << ", L2=" << p.L2 << ", L3=" << p.L3 << " (Q=S(nn-I/3), S_ref=" << S_ref
                  << ", Kappa set to 0)" << std::endl;
        if (!(p.L1 > 0.0) || !((p.L1 + p.L2) > 0.0)) {
            std::cout << "\n[WARNING] Potentially unstable elastic parameters (L1=" << p.L1 << ", L2=" << p.L2
                      << ", L1+L2=" << (p.L1 + p.L2) << ").\n";
            if (!cfg.get_b("proceed_unstable_elastic", false, "Proceed anyway?")) {
                std::cout << "Aborting this run; choose different K1/K2/K3 or do not use Frank mapping." << std::endl;
                return 1;
            }
        }
    } else {
        p.L1 = cfg.get_d("L1", 0.0, "Enter L1");
        p.L2 = cfg.get_d("L2", 0.0, "Enter L2");
        p.L3 = cfg.get_d("L3", 0.0, "Enter L3");
        if (p.L1 != 0 || p.L2 != 0 || p.L3 != 0) p.kappa = 0.0;
        if ((p.L1 != 0.0 || p.L2 != 0.0) && (!(p.L1 > 0.0) || !((p.L1 + p.L2) > 0.0))) {
            std::cout << "\n[ERROR] Unstable elastic parameters for this model (L1=" << p.L1 << ", L2=" << p.L2
                      << ", L1+L2=" << (p.L1 + p.L2) << ").\n";
            if (!cfg.get_b("proceed_unstable_elastic", false, "Proceed anyway?")) {
                std::cout << "Aborting this run; choose stable L1/L2 or use kappa instead." << std::endl;
                return 1;
            }
        }
    }
    p.radius_index = cfg.get_d("droplet_radius", -1.0, "Droplet radius in voxels (<=0: inscribed)");
    p.center_index[0] = cfg.get_d("droplet_center_x", kNaN);
    p.center_index[1] = cfg.get_d("droplet_center_y", kNaN);
    p.center_index[2] = cfg.get_d("droplet_center_z", kNaN);
    const qsrh::Geometry geom = qsrh::make_geometry(p.Nx, p.Ny, p.Nz, p.dx, p.dy, p.dz, p.radius_index, p.center_index);
    const bool xi_guard = cfg.get_b("xi_guard_enabled", true, "Enable correlation-length guard (abort if xi is under-resolved)?");
    if (!qsrh::correlation_length_guard(geom, p.a, p.b, p.c, T, p.T_star, p.modechoice, p.kappa, p.L1, p.L2, p.L3,
                                        xi_guard)) {
        std::cout << "\nAborting this run due to correlation-length guard." << std::endl;
        std::cout << "Tip: decrease the coarsest spacing, enlarge the smallest droplet semi-axis, or move T closer to T* "
                     "to increase xi."
                  << std::endl;
        return 1;
    }
    p.W = cfg.get_d("W", 0.0, "Enter weak anchoring W (J/m^2)");
    p.gamma = cfg.get_d("gamma", 0.1, "Enter gamma (Pa s)");
    const int maxIterations = cfg.get_i("iterations", 100000, "Enter iterations");
    const int printFreq = std::max(1, cfg.get_i("print_freq", 200, "Enter print freq"));
    const double tolerance = cfg.get_d("tolerance", 1e-2, "Enter tolerance");
    const double defect_eps = cfg.get_d("defect_density_abs_eps", 1e-4);
    const double noise_sigma = cfg.get_d("noise_sigma", 0.0);
    p.noise_sigma = noise_sigma;
    p.seed = uint64_t(cfg.get_ll("noise_seed", random_seed >= 0 ? random_seed : 0));
    const int graph_chunk = cfg.get_i("step_graph_chunk", 0);

    qsrh::RowMasks masks;
    std::string gerr;
    if (!qsrh::build_row_masks(geom, p.W, masks, gerr)) {
        std::cerr << "[Geometry] " << gerr << std::endl;
        return 1;
    }
    const double shell_thickness = qsrh::shell_thickness(geom, masks.shell_count);
    if (p.W > 0.0) {
        // machine-parsed by tools/weak_anchor_mesh_check.py
        std::cout << "\n--- Weak Anchoring Shell ---" << std::endl;
        std::cout << "Shell voxels              = " << masks.shell_count << std::endl;
        std::cout << "Shell volume              = " << double(masks.shell_count) * p.dx * p.dy * p.dz << " m^3" << std::endl;
        std::cout << "Estimated surface area    = " << qsrh::surface_area(geom) << " m^2" << std::endl;
        std::cout << "Nominal shell thickness   = " << 2.0 * geom.shell_half << " m" << std::endl;
        std::cout << "Effective shell thickness = " << shell_thickness << " m" << std::endl;
        std::cout << "Weak anchoring scale W/δ  = " << (p.W / shell_thickness) << " J/m^3" << std::endl;
    }

    std::cout << "\nSelect simulation mode: [1] Single Temp, [2] Temp Range, [3] Quench (time-dependent T): ";
    const int sim_mode = cfg.get_i("sim_mode", 1, "");
    const std::string bom_raw = cfg.get_s("boundary_order_mode", "equilibrium", "Boundary order policy [equilibrium/custom]");
    const std::string bom = qsrh::lower(qsrh::trim(bom_raw));
    const bool custom = (bom == "custom" || bom == "fixed");
    if (!bom.empty() && bom != "equilibrium" && !custom)
        std::cout << "[Boundary] Unrecognized boundary_order_mode='" << bom_raw << "'; using equilibrium." << std::endl;
    double boundary_S = 0.0;
    if (custom) {
        const double def = qsrh::shell_order(false, 0.0, p.a, p.b, p.c, T, p.T_star, p.modechoice);
        boundary_S = cfg.get_d("boundary_S", def, "Boundary shell order S_shell for custom policy");
        if (!std::isfinite(boundary_S)) boundary_S = def;
    }
    p.boundary_custom = custom ? 1 : 0;
    p.boundary_S = boundary_S;
    auto S_shell_at = [&](double Tv) {
        return qsrh::shell_order(custom, boundary_S, p.a, p.b, p.c, Tv, p.T_star, p.modechoice);
    };
    if (custom)
        std::cout << "[Boundary] Shell order policy: custom S_shell = " << boundary_S
                  << " (independent of init S0 and temperature)." << std::endl;
    else
        std::cout << "[Boundary] Shell order policy: equilibrium S_shell = max(S_eq(T), 0)." << std::endl;

    Run run;
    run.n = size_t(p.Nx) * p.Ny * p.Nz;
    const size_t n = run.n;
    auto make_ctx = [&](qsr_params& pp) -> int {
        run.p = pp;
        return check(qsr_create(&pp, 0, &run.ctx), "qsr_create");
    };
    // warm start / resume from a Q-tensor file (extension key; overrides init_mode)
    const std::string warm = cfg.kv.count("init_qtensor_file") ? qsrh::trim(cfg.kv.at("init_qtensor_file")) : "";
    auto initial_field = [&](double noise_amp, std::vector<double>& Q) -> int {
        if (!warm.empty()) {
            uint64_t rows = 0;
            std::string err;
            if (!qsrh::load_qtensor_text(warm, p.Nx, p.Ny, p.Nz, Q, rows, err)) {
                std::cerr << "[Init] cannot warm start from " << warm << ": " << err << std::endl;
                return 1;
            }
            std::cout << "[Init] Warm start from " << warm << " (" << rows << " rows)" << std::endl;
            return 0;
        }
        if (mode == 2) qsrh::init_isotropic(Q, geom, noise_amp, random_seed);
        else qsrh::init_director(Q, geom, S0, mode, random_seed);
        return 0;
    };
    auto init_and_upload = [&](double noise_amp) -> int {
        std::vector<double> Q;
        if (initial_field(noise_amp, Q)) return 1;
        return check(qsr_upload_q(run.ctx, 0, Q.data(), n), "upload");
    };
    auto dt_window = [&](double Ta, double Tb, bool semi, double L_stab, int J) {
        qsrh::DtInput in;
        in.dx = p.dx; in.dy = p.dy; in.dz = p.dz; in.gamma = p.gamma;
        in.a = p.a; in.b = p.b; in.c = p.c; in.T_star = p.T_star; in.modechoice = p.modechoice;
        in.kappa = p.kappa; in.L1 = p.L1; in.L2 = p.L2; in.L3 = p.L3;
        in.S_ref_used = S_ref_used; in.S0 = S0; in.T_a = Ta; in.T_b = Tb;
        in.semi_implicit = semi; in.L_stab = L_stab; in.jacobi_iters = J;
        return qsrh::stable_dt(in);
    };
    auto step = [&](int nsteps) -> int {
        if (nsteps <= 0) return 0;
        if (graph_chunk > 1) return check(qsr_step_graph(run.ctx, nsteps, graph_chunk & ~1), "step");
        return check(qsr_step(run.ctx, nsteps), "step");
    };
    qsr_obs_opts oo{};

    if (sim_mode == 3) {
        // ---------------- quench (QSR.cu:2412-2983) ----------------
        const std::string out_dir_name = cfg.get_s("out_dir", "output_quench", "Output directory for quench results");
        const fs::path out_dir = out_dir_name;
        if (fs::exists(out_dir) && cfg.get_b("overwrite_out_dir", true)) fs::remove_all(out_dir);
        fs::create_directories(out_dir);
        const int protocol = cfg.get_i("protocol", 2, "Quench protocol (1=step, 2=ramp)");
        const double T_high = cfg.get_d("T_high", T, "T_high (K)");
        const double T_low = cfg.get_d("T_low", T_high - 5.0, "T_low (K)");
        const int pre = cfg.get_i("pre_equil_iters", 0, "Pre-equilibration iterations at T_high");
        int ramp = 0;
        if (protocol == 2) ramp = std::max(1, cfg.get_i("ramp_iters", 1000, "Ramp iterations (>=1)"));
        const int total = std::max(1, cfg.get_i("total_iters", maxIterations, "Total iterations"));
        const int logFreq = std::max(1, cfg.get_i("logFreq", printFreq, "Log/print freq"));
        const int snapshot_mode = std::clamp(cfg.get_i("snapshot_mode", 2, "Select snapshot mode"), 0, 2);
        int snapshotFreq = 0;
        if (snapshot_mode == 1) snapshotFreq = std::max(0, cfg.get_i("snapshotFreq", 10000));
        double Tc_KZ = kNaN, Tc_win = 0.0;
        int kzFreq = 0, kzAfter = 0, kzExtra = 0;
        bool kz_stop_early = false, save_q_snaps = false;
        if (snapshot_mode == 2) {
            Tc_KZ = cfg.get_d("Tc_KZ", 310.2);
            if (protocol == 2) {
                Tc_win = std::abs(cfg.get_d("Tc_window_K", 0.5));
                kzFreq = std::max(1, cfg.get_i("kzSnapshotFreq", 1000));
            } else {
                kzAfter = std::max(1, cfg.get_i("kzItersAfterStep", 200000));
                kzFreq = std::max(1, cfg.get_i("kzSnapshotFreq", 1000));
            }
            kz_stop_early = cfg.get_b("kz_stop_early", true);
            if (kz_stop_early) kzExtra = std::max(0, cfg.get_i("kzExtraIters", 0));
            save_q_snaps = cfg.get_b("save_qtensor_snapshots", false);
        }
        const double noise_amp = (mode == 2) ? cfg.get_d("noise_amplitude", 1e-3, "Noise amplitude for isotropic init") : 0.0;
        const bool semi = cfg.get_b("use_semi_implicit", true, "Quench: use semi-implicit elastic stabilizer?");
        const bool ldg = (p.L1 != 0.0 || p.L2 != 0.0 || p.L3 != 0.0);
        const double L_base_abs = std::abs(ldg ? p.L1 : p.kappa);
        double L_stab = 0.0;
        int J = 0;
        if (semi) {
            L_stab = cfg.get_d("L_stab", L_base_abs > 0.0 ? L_base_abs : 0.0);
            if (!std::isfinite(L_stab) || L_stab < 0.0) L_stab = 0.0;
            if (L_base_abs > 0.0 && L_stab > L_base_abs) {
                std::cout << "[SemiImplicit] Clamping L_stab from " << L_stab << " to |L_base|=" << L_base_abs
                          << " to avoid anti-diffusive explicit remainder." << std::endl;
                L_stab = L_base_abs;
            }
            J = std::max(0, cfg.get_i("jacobi_iters", 25));
        }
        const qsrh::DtWindow w = dt_window(T_high, T_low, semi, L_stab, J);
        print_dt(w);
        double dt = cfg.get_d("dt", w.dt_max, "Enter time step dt (s)");
        if (dt > w.dt_max) dt = w.dt_max;
        const bool guard = cfg.get_b("enable_instability_guard", true);
        if (cfg.kv.count("enable_adaptive_dt"))
            std::cout << "[Protocol] Ignoring legacy enable_adaptive_dt=" << cfg.kv.at("enable_adaptive_dt")
                      << " in quench mode. Kibble-Zurek runs use fixed dt." << std::endl;
        std::cout << "[Protocol] Quench/KZ runs enforce fixed dt = " << dt
                  << " s. Instability handling may abort the run, but it will not change dt." << std::endl;
        const double S_abort = guard ? cfg.get_d("S_abort", 2.0) : 0.0;
        const double checker_abort = guard ? cfg.get_d("checker_rel_abort", 0.10) : 0.0;
        const bool limiter = cfg.get_b("enable_q_limiter", false);
        const double S_cap = limiter ? cfg.get_d("S_cap", 1.2) : 0.0;
        const bool early = cfg.get_b("enable_early_stop", false);
        const double rad_thr = early ? cfg.get_d("radiality_threshold", 0.998) : 0.0;
        const bool console = cfg.get_b("console_output", true);
        const bool log_def = cfg.get_b("log_defects_2d", true);
        const bool track_def = log_def || early;
        int zdef = p.Nz / 2;
        double def_thr = 0.1, def_cut = 0.25;
        if (track_def) {
            zdef = std::clamp(cfg.get_i("defects_z_slice", zdef), 0, p.Nz - 1);
            def_thr = cfg.get_d("defects_S_threshold", def_thr);
            def_cut = cfg.get_d("defects_charge_cutoff", def_cut);
        }
        if (early) std::cout << "[Convergence] Quench early-stop uses energy + 2D defect proxy stability; radiality is diagnostic only." << std::endl;
        const bool log_xi = cfg.get_b("log_xi_grad_2d", true);
        int zxi = zdef;
        double xi_thr = def_thr;
        if (log_xi) {
            zxi = std::clamp(cfg.get_i("xi_z_slice", zxi), 0, p.Nz - 1);
            xi_thr = cfg.get_d("xi_S_threshold", xi_thr);
        }
        p.dt = dt;
        p.q_norm_cap = (limiter && S_cap > 0.0) ? S_cap * std::sqrt(2.0 / 3.0) : 0.0;
        p.protocol = protocol == 1 ? QSR_PROTOCOL_STEP : QSR_PROTOCOL_RAMP;
        p.T_high = T_high;
        p.T_low = T_low;
        p.pre_equil_iters = pre;
        p.ramp_iters = ramp;
        if (make_ctx(p) || init_and_upload(noise_amp)) return 1;
        std::ofstream qlog((out_dir / "quench_log.dat").string());
        qlog << "iteration,time_s,dt_s,T_K,bulk,elastic,anchoring,total,radiality,avg_S,max_S,checker_rel,"
                "defect_density_per_plaquette,defect_plaquettes_used,xi_grad_proxy,xi_grad_edges_used\n";
        auto T_at = [&](int it) {
            if (it < pre) return T_high;
            if (protocol == 1) return T_low;
            const int t = it - pre;
            if (t <= 0) return T_high;
            if (t >= ramp) return T_low;
            return T_high + (T_low - T_high) * (double(t) / double(ramp));
        };
        auto reached_final = [&](int it) { return it >= pre && (protocol == 1 || it >= pre + ramp); };
        if (!custom)
            std::cout << "[Boundary] Quench shell order follows S_eq(T): S_shell(T_high)=" << S_shell_at(T_high)
                      << ", S_shell(T_low)=" << S_shell_at(T_low) << std::endl;
        else
            std::cout << "[Boundary] Quench shell order fixed at S_shell=" << boundary_S << std::endl;
        double t_phys = 0.0, prev_F = kNaN, prev_def = kNaN;
        bool entered = false, aborted = false;
        int exit_iter = -1, pending = 0;
        oo.z_slice = zdef; oo.S_threshold = def_thr; oo.charge_cutoff = def_cut;
        oo.xi_z_slice = zxi; oo.xi_S_threshold = xi_thr;
        for (int it = 0; it < total; ++it) {
            const double Tc = T_at(it);
            ++pending;
            t_phys += dt;
            const bool do_log = (it % logFreq == 0) || (it == total - 1);
            bool in_win = false, do_snap = false;
            if (snapshot_mode == 1) {
                do_snap = snapshotFreq > 0 && ((it % snapshotFreq == 0) || it == total - 1);
            } else if (snapshot_mode == 2) {
                if (protocol == 2) {
                    in_win = (std::isfinite(Tc_KZ) && Tc_win > 0.0) ? std::abs(Tc - Tc_KZ) <= Tc_win
                                                                     : (std::isfinite(Tc_KZ) && std::abs(Tc - Tc_KZ) <= 1e-12);
                } else {
                    in_win = it >= pre && it <= pre + kzAfter;
                }
                const bool in_post = kz_stop_early && kzExtra > 0 && exit_iter >= 0 && it > exit_iter && it <= exit_iter + kzExtra;
                do_snap = kzFreq > 0 && (it % kzFreq == 0) && (in_win || in_post);
            }
            bool stop = false;
            if (do_log || do_snap) {
                if (semi && L_stab > 0.0 && J > 0) {
                    if (check(qsr_step_imex(run.ctx, pending, L_stab, J), "step_imex")) return 1;
                } else if (step(pending)) {
                    return 1;
                }
                pending = 0;
                oo.want_defects = (do_log && track_def) ? 1 : 0;
                oo.want_xi = (do_log && log_xi) ? 1 : 0;
                qsr_obs o{};
                if (check(qsr_observables(run.ctx, 0, &oo, &o), "observables")) return 1;
                if (do_log) {
                    qlog << it << "," << t_phys << "," << dt << "," << Tc << "," << o.bulk << "," << o.elastic << ","
                         << o.anchoring << "," << o.total << "," << o.radiality << "," << o.avg_S << "," << o.max_S << ","
                         << o.checker_rel << "," << o.defect_density << "," << o.plaquettes << "," << o.xi_grad_proxy
                         << "," << o.xi_edges << "\n";
                    qlog.flush();
                    if (console)
                        std::cout << "Iter " << it << "  t=" << t_phys << " s  T=" << Tc << " K  F=" << o.total
                                  << " (bulk=" << o.bulk << ", el=" << o.elastic << ", anch=" << o.anchoring
                                  << ")  Rbar=" << o.radiality << "  <S>=" << o.avg_S << "  maxS=" << o.max_S
                                  << "  checker=" << o.checker_rel << std::endl;
                    if (guard) {
                        const bool s_bad = !std::isfinite(o.max_S) || (S_abort > 0.0 && o.max_S > S_abort);
                        const bool cb_bad = std::isfinite(o.checker_rel) && checker_abort > 0.0 && o.checker_rel > checker_abort;
                        if (s_bad || cb_bad) {
                            std::cerr << "\n[InstabilityGuard] Detected runaway/aliasing at iter " << it
                                      << ": max_S=" << o.max_S << ", checker_rel=" << o.checker_rel << ". dt=" << dt
                                      << " s" << std::endl;
                            std::cerr << "[InstabilityGuard] Quench/KZ contract requires fixed dt. Aborting instead of"
                                         " changing dt. Try smaller dt, stronger stabilization, or milder elastic settings."
                                      << std::endl;
                            aborted = true;
                            break;
                        }
                    }
                    if (early && reached_final(it)) {
                        bool e_conv = false;
                        if (std::isfinite(prev_F) && prev_F != 0.0) e_conv = std::abs((o.total - prev_F) / prev_F) < tolerance;
                        const bool d_conv = o.plaquettes > 0 && std::isfinite(prev_def) && std::isfinite(o.defect_density) &&
                                            std::abs(o.defect_density - prev_def) < defect_eps;
                        if (e_conv && d_conv) {
                            std::cout << "\n=== Early-stop: converged at final T (energy + defect proxy stable; Rbar="
                                      << o.radiality << ", defect2D=" << o.defect_density << ", radiality_threshold_met="
                                      << ((rad_thr <= 0.0 || o.radiality > rad_thr) ? "yes" : "no") << ") ===\n";
                            stop = true;
                        }
                        if (!stop) {
                            prev_F = o.total;
                            prev_def = o.plaquettes > 0 ? o.defect_density : kNaN;
                        }
                    }
                }
                if (!stop && do_snap) {
                    if (run.save_nematic_only((out_dir / ("nematic_field_iter_" + std::to_string(it) + ".dat")).string())) return 1;
                    if (save_q_snaps && run.save_q_only((out_dir / ("Qtensor_output_iter_" + std::to_string(it) + ".dat")).string()))
                        return 1;
                }
            }
            if (stop) break;
            if (snapshot_mode == 2 && kz_stop_early) {
                bool kz_stop = false;
                if (protocol == 2) {
                    if (in_win) entered = true;
                    const bool below = std::isfinite(Tc_KZ) && Tc < (Tc_KZ - Tc_win);
                    if (entered && !in_win && below) {
                        if (exit_iter < 0) exit_iter = it;
                        if (it >= exit_iter + kzExtra) {
                            std::cout << "\n=== KZ mode: stopping after leaving Tc window ===\n";
                            kz_stop = true;
                        }
                    }
                } else if (it >= pre + kzAfter + kzExtra) {
                    std::cout << "\n=== KZ mode: stopping after post-step recording window ===\n";
                    kz_stop = true;
                }
                if (kz_stop) {
                    if (pending > 0) {
                        if (semi && L_stab > 0.0 && J > 0) {
                            if (check(qsr_step_imex(run.ctx, pending, L_stab, J), "step_imex")) return 1;
                        } else if (step(pending)) {
                            return 1;
                        }
                        pending = 0;
                    }
                    break;
                }
            }
        }
        if (aborted) {
            std::cerr << "[Quench] Run aborted by the fixed-dt instability guard. Final state was not saved." << std::endl;
            return 1;
        }
        if (pending > 0) {
            if (semi && L_stab > 0.0 && J > 0) {
                if (check(qsr_step_imex(run.ctx, pending, L_stab, J), "step_imex")) return 1;
            } else if (step(pending)) {
                return 1;
            }
        }
        if (run.save_final((out_dir / "nematic_field_final.dat").string(), (out_dir / "Qtensor_output_final.dat").string()))
            return 1;
    } else if (sim_mode == 2) {
        // ---------------- temperature sweep (QSR.cu:2985-3124) ----------------
        if (fs::exists("output_temp_sweep")) fs::remove_all("output_temp_sweep");
        fs::create_directory("output_temp_sweep");
        std::ofstream sweep("output_temp_sweep/summary.dat");
        sweep << "temperature,final_energy_including_anchoring,average_S\n";
        const double T_start = cfg.get_d("T_start", 295.0, "Start T");
        const double T_end = cfg.get_d("T_end", 315.0, "End T");
        double T_step = cfg.get_d("T_step", 1.0, "Step T");
        if (T_step == 0.0) {
            std::cout << "Step T cannot be 0. Aborting sweep." << std::endl;
            return 0;
        }
        if ((T_end - T_start) * T_step < 0.0) T_step = -T_step;
        {
            const qsrh::BulkABC k0 = qsrh::bulk_abc(p.a, p.b, p.c, 0.0, 0.0, p.modechoice);
            const double T_NI = qsrh::t_from_a(qsrh::a_coexistence(k0.B, k0.C), p.a, p.T_star);
            const double T_sp = qsrh::t_from_a(qsrh::a_spinodal_nematic(k0.B, k0.C), p.a, p.T_star);
            if (std::isfinite(T_NI) && std::max(T_start, T_end) < T_NI)
                std::cout << "\n[NOTE] Sweep max(T) < T_NI (" << T_NI
                          << " K). You should not expect a bulk I->N transition within this range." << std::endl;
            if (std::isfinite(T_sp) && std::min(T_start, T_end) > T_sp)
                std::cout << "\n[NOTE] Sweep min(T) > T_spin,N (" << T_sp
                          << " K). Bulk nematic minimum does not exist; expect isotropic." << std::endl;
        }
        const double noise_amp = (mode == 2) ? cfg.get_d("noise_amplitude", 1e-3) : 0.0;
        std::vector<double> Q;
        if (initial_field(noise_amp, Q)) return 1;
        const qsrh::DtWindow w = dt_window(T_start, T_end, false, 0.0, 0);
        print_dt(w);
        double dt = cfg.get_d("dt", w.dt_max, "Enter time step dt (s)");
        if (dt > w.dt_max) dt = w.dt_max;
        const double tau = geom.min_radius * geom.min_radius / w.D_total;
        const double t_min = 0.25 * tau;
        std::cout << "Estimated alignment time tau_align ~ " << tau << " s, enforcing minimum sweep time ~ " << t_min
                  << " s" << std::endl;
        p.dt = dt;
        p.q_norm_cap = 0.0;
        p.protocol = QSR_PROTOCOL_CONSTANT;
        auto done = [&](double Tv) { return T_step > 0.0 ? (Tv > T_end + 1e-12) : (Tv < T_end - 1e-12); };
        oo.z_slice = -1; oo.S_threshold = 0.1; oo.charge_cutoff = 0.25; oo.xi_z_slice = -1; oo.xi_S_threshold = -1;
        oo.want_defects = 0; oo.want_xi = 0;
        // One context per temperature (the shell order and bulk A are per-context constants);
        // the field carries over from the previous temperature (QSR.cu:3054-3122).
        for (double Tc = T_start; !done(Tc); Tc += T_step) {
            std::cout << "\n--- Running simulation for T = " << Tc << " K with S_shell = " << S_shell_at(Tc) << " ("
                      << (custom ? "custom" : "equilibrium") << ") ---\n";
            const std::string tdir = "output_temp_sweep/T_" + std::to_string(Tc);
            fs::create_directory(tdir);
            qsr_params pt = p;
            pt.T_high = Tc;
            if (make_ctx(pt)) return 1;
            if (check(qsr_upload_q(run.ctx, 0, Q.data(), n), "upload")) return 1;
            double prev_F = std::numeric_limits<double>::max(), t_phys = 0.0;
            int pending = 0;
            for (int it = 0; it < maxIterations; ++it) {
                ++pending;
                t_phys += dt;
                if (it > 0 && it % printFreq == 0) {
                    if (step(pending)) return 1;
                    pending = 0;
                    qsr_obs o{};
                    if (check(qsr_observables(run.ctx, 0, &oo, &o), "observables")) return 1;
                    std::cout << "  Iter " << it << ", Free Energy: " << o.total << std::endl;
                    if (t_phys >= t_min && prev_F != 0.0 &&
                        (std::abs((o.total - prev_F) / prev_F) < tolerance || o.total == 0.0)) {
                        std::cout << "  Convergence reached at iteration " << it << std::endl;
                        break;
                    }
                    prev_F = o.total;
                }
            }
            if (step(pending)) return 1;
            qsr_obs o{};
            if (check(qsr_observables(run.ctx, 0, &oo, &o), "observables")) return 1;
            sweep << Tc << "," << o.total << "," << o.avg_S << "\n";
            sweep.flush();
            if (run.save_final(tdir + "/nematic_field_final.dat", tdir + "/Qtensor_output_final.dat")) return 1;
            if (check(qsr_download_q(run.ctx, 0, Q.data(), n), "download")) return 1;
            std::cout << "  Saved final state for T = " << Tc << " K. Average S = " << o.avg_S << std::endl;
            qsr_destroy(run.ctx);
            run.ctx = nullptr;
        }
    } else {
        // ---------------- single temperature (QSR.cu:3126-3476) ----------------
        int submode = cfg.get_i("submode", 1, "Select submode: [1] full energy, [2] energy components");
        if (submode != 1 && submode != 2) submode = 1;
        std::cout << "Submode selected: " << (submode == 1 ? "full energy" : "energy components") << std::endl;
        if (fs::exists("output"))
            for (const auto& e : fs::directory_iterator("output")) fs::remove_all(e.path());
        fs::create_directory("output");
        const qsrh::DtWindow w = dt_window(T, T, false, 0.0, 0);
        const double tau = geom.min_radius * geom.min_radius / w.D_total;
        std::cout << "\n--- Time Scale Analysis ---\nDiffusion coefficient D = " << w.D_total
                  << " m²/s\nDroplet radius R ~ " << geom.min_radius << " m\nEstimated alignment time tau_align ~ " << tau
                  << " s" << std::endl;
        print_dt(w);
        double dt = cfg.get_d("dt", w.dt_max, "Enter time step dt (s)");
        if (dt > w.dt_max) dt = w.dt_max;
        const double noise_amp = (mode == 2) ? cfg.get_d("noise_amplitude", 1e-3) : 0.0;
        p.dt = dt;
        p.q_norm_cap = 0.0;
        p.protocol = QSR_PROTOCOL_CONSTANT;
        p.T_high = T;
        if (make_ctx(p) || init_and_upload(noise_amp)) return 1;
        const double S_shell = S_shell_at(T);
        std::cout << "[Boundary] Single-temp shell order at T=" << T << " K: S_shell=" << S_shell << " ("
                  << (custom ? "custom" : "equilibrium") << " policy)." << std::endl;
        const bool console = cfg.get_b("console_output", true);
        const int zdef = std::clamp(cfg.get_i("defects_z_slice", p.Nz / 2), 0, p.Nz - 1);
        const double def_thr = cfg.get_d("defects_S_threshold", 0.1), def_cut = cfg.get_d("defects_charge_cutoff", 0.25);
        const bool log_xi = cfg.get_b("log_xi_grad_2d", true);
        int zxi = zdef;
        double xi_thr = def_thr;
        if (log_xi) {
            zxi = std::clamp(cfg.get_i("xi_z_slice", zxi), 0, p.Nz - 1);
            xi_thr = cfg.get_d("xi_S_threshold", xi_thr);
        }
        const bool snaps = cfg.get_b("save_iteration_snapshots", true);
        const bool time_gate = cfg.get_b("early_stop_min_alignment_time", true);
        std::cout << "[Convergence] Single-temp stopping uses energy + 2D defect proxy stability; radiality is diagnostic only." << std::endl;
        std::ofstream elog(submode == 1 ? "free_energy_vs_iteration.dat" : "energy_components_vs_iteration.dat");
        elog << "iteration,bulk,elastic,anchoring,total,radiality,time,avg_S,max_S,defect_density_per_plaquette,"
                "defect_plaquettes_used,xi_grad_proxy,xi_grad_edges_used\n";
        oo.z_slice = zdef; oo.S_threshold = def_thr; oo.charge_cutoff = def_cut; oo.xi_z_slice = zxi;
        oo.xi_S_threshold = xi_thr; oo.want_defects = 1; oo.want_xi = log_xi ? 1 : 0;
        double prev_F = std::numeric_limits<double>::max(), prev_def = kNaN, t_phys = 0.0;
        const double t_min = 0.5 * tau;
        int pending = 0;
        for (int it = 0; it < maxIterations; ++it) {
            t_phys += dt;
            ++pending;
            if (it % printFreq != 0) continue;
            if (step(pending)) return 1;
            pending = 0;
            qsr_obs o{};
            if (check(qsr_observables(run.ctx, 0, &oo, &o), "observables")) return 1;
            elog << it << "," << o.bulk << "," << o.elastic << "," << o.anchoring << "," << o.total << "," << o.radiality
                 << "," << t_phys << "," << o.avg_S << "," << o.max_S << "," << o.defect_density << "," << o.plaquettes
                 << "," << o.xi_grad_proxy << "," << o.xi_edges << "\n";
            if (console)
                std::cout << "Iter " << it << "  t=" << t_phys << " s  F=" << o.total << " (bulk=" << o.bulk
                          << ", el=" << o.elastic << ", anch=" << o.anchoring << ")  R̄=" << o.radiality
                          << "  <S>=" << o.avg_S << "  maxS=" << o.max_S << "  defect2D=" << o.defect_density
                          << "  xiGrad=" << o.xi_grad_proxy << std::endl;
            if (snaps && run.save_nematic_only("output/nematic_field_iter_" + std::to_string(it) + ".dat")) return 1;
            if (it > 100 && prev_F != 0.0) {
                const bool e_conv = std::abs((o.total - prev_F) / prev_F) < tolerance;
                const bool t_ok = !time_gate || t_phys > t_min;
                const bool d_conv = o.plaquettes > 0 && std::isfinite(prev_def) && std::isfinite(o.defect_density) &&
                                    std::abs(o.defect_density - prev_def) < defect_eps;
                if (e_conv && d_conv && t_ok) {
                    std::cout << "\n=== Convergence Achieved (energy + defect proxy stable; Rbar=" << o.radiality
                              << ", defect2D=" << o.defect_density << ", radiality_threshold_met="
                              << (o.radiality > 0.998 ? "yes" : "no") << ") ===";
                    break;
                }
            }
            prev_F = o.total;
            prev_def = o.plaquettes > 0 ? o.defect_density : kNaN;
        }
        if (pending > 0 && step(pending)) return 1;
        elog.close();
        if (run.save_final("nematic_field_final.dat", "Qtensor_output_final.dat")) return 1;
    }
    std::cout << "\nSimulation finished." << std::endl;
    return 0;
}
