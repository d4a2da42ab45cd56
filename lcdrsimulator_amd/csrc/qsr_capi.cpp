// qsr_capi.cpp -- the C ABI (include/qsr_amd.h): context lifetime, field
// transfer, the step loop and the log-point observables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/qsr_amd.h"
#include "qsr_ctx.hpp"
#include "qsr_host.hpp"
#include "qsr_kernels.hpp"
#include "qsr_surr.hpp"

// Scratch of qsr_energies_noalloc: the volume partials of one box (the volume
// passes launch at most 4096 blocks of NPART slots), then NPART finalized sums
// and one total per listed box.  One copy per GPU, held until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns.
constexpr int kNoallocMaxBoxes = 4096;
constexpr long long kNoallocPart = 4096LL * qsro::NPART;
constexpr long long kNoallocDoubles = kNoallocPart + (long long)kNoallocMaxBoxes * (qsro::NPART + 1);
static __device__ __align__(256) double scratch[kNoallocDoubles];

__global__ void energy_totals_kernel(const double* sums, int n, double* totals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const double* w = sums + (long long)i * qsro::NPART;
        totals[i] = (w[0] + w[1]) + w[2];  // bulk + elastic + anchoring, as the host sums them
    }
}

namespace {

thread_local std::string g_err;

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU, set at context creation

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(QSR_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Release-path calls whose status cannot be acted on (destructors, cleanup
// after an earlier error) go through this to make the discard explicit.
inline void discard(hipError_t) {}

constexpr double kNaN = std::numeric_limits<double>::quiet_NaN();
constexpr int kDynChunk = 512;  // steps per dynamic-table upload

struct Box {
    qsr_params p{};
    qsrh::Geometry g;
    qsrh::RowMasks masks;
    double thickness = 0, area = 0;
    long long off = 0, nvox = 0, rows_off = 0, nface = 0;
    int64_t iter = 0;  // steps completed
    double last_T = 0, last_S = 0;
    bool uploaded = false;
    bool active = true;  // inactive boxes are held: not stepped, iteration frozen
    double m = 0;  // dt / gamma

    double T_at(int64_t it) const {
        // compute_T_current (QSR.cu:2679-2687)
        if (p.protocol == QSR_PROTOCOL_CONSTANT) return p.T_high;
        if (it < p.pre_equil_iters) return p.T_high;
        if (p.protocol == QSR_PROTOCOL_STEP) return p.T_low;
        const int64_t t = it - p.pre_equil_iters;
        if (t <= 0) return p.T_high;
        if (p.ramp_iters <= 0 || t >= p.ramp_iters) return p.T_low;
        return p.T_high + (p.T_low - p.T_high) * (double(t) / double(p.ramp_iters));
    }
    double S_at(double T) const {
        if (p.periodic) return 0.0;
        return qsrh::shell_order(p.boundary_custom != 0, p.boundary_S, p.a, p.b, p.c, T, p.T_star, p.modechoice);
    }
    qsrk::StepDyn dyn_at(int64_t it) const {
        const double T = T_at(it);
        const qsrh::BulkABC k = qsrh::bulk_abc(p.a, p.b, p.c, T, p.T_star, p.modechoice);
        qsrk::StepDyn d{};
        d.cA = float(m * k.A);
        d.S_shell = float(S_at(T));
        d.iter = unsigned(it);
        return d;
    }
};

struct Ctx {
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<Box> boxes;
    long long cs = 0;  // floats per component plane
    float* q[2] = {nullptr, nullptr};
    int cur = 0;
    bool full = false, noise = false;
    int elastic = 0;  // union of the boxes' anisotropic terms: 1 = L2, 2 = L3
    int ch = 32;
    int variant = 0;  // tile shape (qsrk::tile_k / tile_j), chosen by qsr_autotune
    std::vector<qsrk::BoxDev> h_boxes;
    qsrk::BoxDev* d_boxes = nullptr;
    std::vector<int4> h_tiles;
    int4* d_tiles = nullptr;
    int ntiles = 0, tiles_cap = 0;
    // two-step L1 kernel: its own tile table (X2_OUT_K x X2_OUT_J outputs, ch2 planes)
    bool fuse = true;      // pair steps when the context allows it (qsr_set_step_fusion)
    int ch2 = 16;
    int4* d_tiles2 = nullptr;
    int ntiles2 = 0, tiles2_cap = 0;
    long long tile_voxels2 = 0;
    long long tile_voxels = 0, active_voxels = 0, box_voxels = 0;
    qsrk::RowMask* d_step_rows = nullptr;
    qsrk::RowMask* d_obs_rows = nullptr;
    int4* d_face_list = nullptr;  // (box, i, j, k) of the box-face voxels that can change
    long long nface = 0;
    qsrk::StepDyn* d_dyn = nullptr;
    qsrk::StepDyn* h_dyn = nullptr;  // pinned
    double* d_part = nullptr;
    long long part_cap = 0;
    double* d_out = nullptr;
    double* d_nf_slice[2] = {nullptr, nullptr};
    long long nf_slice_cap = 0;
    double* d_nf3 = nullptr;
    long long nf3_cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    double last_ms = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_chunk = 0, graph_cur = -1;
    // semi-implicit path
    float* q_rhs = nullptr;
    float4* d_imex = nullptr;   // per box alpha/d^2 (RHS mode of the fused kernels)
    float4* d_jcoef = nullptr;  // per box alpha/d^2 and the Jacobi denominator
    unsigned* d_jflags = nullptr;  // sweep convergence flags, 2 rows x jflag_n
    int jflag_n = 0, jflag_row = 0;
    long long* d_vox_prefix = nullptr;
    long long nvox_total = 0;
    float4* d_rhs0 = nullptr;  // per box (0, 0, 0, 1): the RHS mode with no stabiliser (residuals)
    // many-box reductions (energies, residuals): per-box partial regions and result slots
    double* d_mpart = nullptr;
    long long mpart_cap = 0;
    double* d_mout = nullptr;
    long long mout_cap = 0;
    // their per-call tables: the boxes' observable parameters, block prefix and box list
    qsro::ObsParams* d_mparams = nullptr;
    long long* d_mprefix = nullptr;
    int* d_mlist = nullptr;
    int mtab_cap = 0;

    ~Ctx() {
        if (gexec) discard(hipGraphExecDestroy(gexec));
        if (graph) discard(hipGraphDestroy(graph));
        for (float* p : q) discard(hipFree(p));
        for (void* p : {(void*)d_boxes, (void*)d_tiles, (void*)d_tiles2, (void*)d_step_rows, (void*)d_obs_rows, (void*)d_face_list,
                        (void*)d_dyn, (void*)d_part, (void*)d_out, (void*)d_nf_slice[0], (void*)d_nf_slice[1],
                        (void*)d_nf3, (void*)q_rhs, (void*)d_imex, (void*)d_jcoef, (void*)d_vox_prefix,
                        (void*)d_jflags, (void*)d_rhs0, (void*)d_mpart, (void*)d_mout, (void*)d_mparams,
                        (void*)d_mprefix, (void*)d_mlist})
            discard(hipFree(p));
        if (h_dyn) discard(hipHostFree(h_dyn));
        for (hipEvent_t e : ev)
            if (e) discard(hipEventDestroy(e));
        if (st) discard(hipStreamDestroy(st));
    }
};

int check_params(const qsr_params& p, int idx) {
    auto bad = [&](const char* what) {
        return fail(QSR_E_ARG, "box " + std::to_string(idx) + ": " + what);
    };
    if (p.Nx < 3 || p.Ny < 3 || p.Nz < 3) return bad("Nx, Ny, Nz must be >= 3");
    if (p.Nz > 32767) return bad("Nz > 32767 not supported");
    if (!(p.dx > 0 && p.dy > 0 && p.dz > 0)) return bad("spacings must be positive");
    if (!(p.gamma > 0)) return bad("gamma must be positive");
    if (!(p.dt > 0) || !std::isfinite(p.dt)) return bad("dt must be positive");
    if (p.protocol < 0 || p.protocol > 2) return bad("unknown protocol");
    if (p.noise_sigma < 0) return bad("noise_sigma must be >= 0");
    return QSR_OK;
}

qsrk::BoxDev make_boxdev(const Box& b) {
    const qsr_params& p = b.p;
    qsrk::BoxDev d{};
    d.Nx = p.Nx; d.Ny = p.Ny; d.Nz = p.Nz;
    d.off = b.off;
    d.rows_off = b.rows_off;
    const double m = b.m;
    const qsrh::BulkABC k = qsrh::bulk_abc(p.a, p.b, p.c, p.T_high, p.T_star, p.modechoice);
    d.cB = float(m * k.B);
    d.cC = float(m * k.C);
    const bool ldg = (p.L1 != 0.0 || p.L2 != 0.0 || p.L3 != 0.0);
    const double Lk = ldg ? p.L1 : p.kappa;
    d.lx = float(m * Lk / (p.dx * p.dx));
    d.ly = float(m * Lk / (p.dy * p.dy));
    d.lz = float(m * Lk / (p.dz * p.dz));
    int fl = 0;
    if (p.periodic) {
        fl |= qsrk::F_PERIODIC;
    } else {
        if (p.W != 0.0 && b.thickness > 0.0) {
            fl |= qsrk::F_ANCHOR;
            d.cW = float(m * p.W / b.thickness);
        }
        if (!(p.W > 0.0)) fl |= qsrk::F_BC;
        if (p.W == 0.0) fl |= qsrk::F_SKIPSHELL;
    }
    if (p.noise_sigma > 0.0) fl |= qsrk::F_NOISE;
    if (p.q_norm_cap > 0.0) {
        fl |= qsrk::F_CAP;
        d.cap2 = float(p.q_norm_cap * p.q_norm_cap);
    }
    if (p.L2 != 0.0) fl |= qsrk::F_L2;
    if (p.L3 != 0.0) fl |= qsrk::F_L3;
    d.flags = fl;
    d.sigma = float(p.noise_sigma);
    d.ci = float(b.g.cx / p.dx);
    d.cj = float(b.g.cy / p.dy);
    d.ck = float(b.g.cz / p.dz);
    d.sx = float(p.dx / (b.g.rx * b.g.rx));
    d.sy = float(p.dy / (b.g.ry * b.g.ry));
    d.sz = float(p.dz / (b.g.rz * b.g.rz));
    d.hx = float(1.0 / (2.0 * p.dx)); d.hy = float(1.0 / (2.0 * p.dy)); d.hz = float(1.0 / (2.0 * p.dz));
    d.fx = float(1.0 / p.dx); d.fy = float(1.0 / p.dy); d.fz = float(1.0 / p.dz);
    d.ix2 = float(1.0 / (p.dx * p.dx)); d.iy2 = float(1.0 / (p.dy * p.dy)); d.iz2 = float(1.0 / (p.dz * p.dz));
    d.i4xy = float(1.0 / (4.0 * p.dx * p.dy));
    d.i4xz = float(1.0 / (4.0 * p.dx * p.dz));
    d.i4yz = float(1.0 / (4.0 * p.dy * p.dz));
    d.mL2 = float(m * p.L2);
    d.mL3 = float(m * p.L3);
    d.seed = p.seed;
    return d;
}

// Does any interior output voxel of the tile change in a step?
// Smallest and largest k of the voxels a step changes (or re-imposes) in rows
// i in [ib, ie), j in [jb, je), within the interior k range [1, Nz-2]; false if
// there are none.
bool live_k_bounds(const Box& b, int ib, int ie, int jb, int je, int& lo, int& hi) {
    const qsr_params& p = b.p;
    const bool bc = !(p.W > 0.0);
    lo = p.Nz;
    hi = -1;
    for (int i = ib; i < ie; ++i)
        for (int j = jb; j < je; ++j) {
            const int16_t* s = &b.masks.step[(size_t(i) * p.Ny + j) * 8];
            int a = std::max<int>(1, s[0]), e = std::min<int>(p.Nz - 2, s[1]);
            if (a <= e) {
                lo = std::min(lo, a);
                hi = std::max(hi, e);
            }
            if (bc) {  // shell voxels outside the inner interval [s4, s5]
                const int o0 = std::max<int>(1, s[2]), o1 = std::min<int>(p.Nz - 2, s[3]);
                if (o0 > o1) continue;
                const bool inner = s[4] <= s[5];
                const int first = (inner && o0 >= s[4] && o0 <= s[5]) ? s[5] + 1 : o0;
                const int last = (inner && o1 >= s[4] && o1 <= s[5]) ? s[4] - 1 : o1;
                if (first <= o1) lo = std::min(lo, first);
                if (last >= o0) hi = std::max(hi, last);
            }
        }
    return lo <= hi;
}

// Tiles of tk x tj outputs (k x j) and ch planes over the active boxes, in XCD-aware launch
// order, uploaded to *d (ntiles entries, padding included); voxels = outputs covered.
int make_tiles(Ctx& c, int tk, int tj, int ch, int4** d, int* cap, int* ntiles, long long* voxels,
               std::vector<int4>* host) {
    std::vector<int4> nat;
    *voxels = 0;
    for (size_t bi = 0; bi < c.boxes.size(); ++bi) {
        const Box& b = c.boxes[bi];
        if (!b.active) continue;
        const int Nx = b.p.Nx, Ny = b.p.Ny, Nz = b.p.Nz;
        for (int i0 = 0; i0 < Nx; i0 += ch)
            for (int j0 = 0; j0 < Ny; j0 += tj) {
                const long long ij = (long long)std::min(ch, Nx - i0) * std::min(tj, Ny - j0);
                if (b.p.periodic) {  // every voxel steps: tiles on the k grid
                    for (int k0 = 0; k0 < Nz; k0 += tk) {
                        nat.push_back(make_int4(int(bi), i0, j0, k0));
                        *voxels += ij * std::min(tk, Nz - k0);
                    }
                    continue;
                }
                // droplet: cover only the k range the rows of this (slab, block) can change,
                // with the leftover of the last tile split over both ends (the kernels clamp
                // loads at the box faces and write interior voxels only)
                int lo, hi;
                if (!live_k_bounds(b, std::max(i0, 1), std::min(i0 + ch, Nx - 1), std::max(j0, 1),
                                   std::min(j0 + tj, Ny - 1), lo, hi))
                    continue;
                const int span = hi - lo + 1, n = (span + tk - 1) / tk;
                int start = lo - (n * tk - span) / 2;
                start = n * tk >= Nz ? 0 : std::min(std::max(start, 0), Nz - n * tk);
                for (int t = 0; t < n; ++t) {
                    const int k0 = start + t * tk;
                    if (k0 >= Nz) break;
                    nat.push_back(make_int4(int(bi), i0, j0, k0));
                    *voxels += ij * std::min(tk, Nz - k0);
                }
            }
    }
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
    // so launch position 8*r + x carries tile x*q + r: each XCD (and its L2)
    // sweeps one contiguous run of neighbouring tiles and shares their halos.
    const int T = int(nat.size());
    const int q = (T + 7) / 8;
    std::vector<int4> order(size_t(8) * q, make_int4(-1, 0, 0, 0));
    for (int x = 0; x < 8; ++x)
        for (int r = 0; r < q; ++r) {
            const int n = x * q + r;
            if (n < T) order[size_t(8) * r + x] = nat[n];
        }
    *ntiles = int(order.size());
    if (*ntiles > *cap) {  // the table keeps its allocation while the new one fits
        discard(hipFree(*d));
        *d = nullptr;
        *cap = 0;
        HIP_TRY(hipMalloc(d, sizeof(int4) * size_t(*ntiles)));
        *cap = *ntiles;
    }
    if (*ntiles > 0) HIP_TRY(hipMemcpy(*d, order.data(), sizeof(int4) * size_t(*ntiles), hipMemcpyHostToDevice));
    if (host) *host = std::move(order);
    return QSR_OK;
}

// Can this context pair its explicit steps in the two-step kernel?  L1 / kappa only, no
// noise, no box face that can change, 32-bit buffer offsets.
bool fusable(const Ctx& c) { return c.fuse && c.elastic == 0 && !c.noise && c.nface == 0 && c.cs < (1LL << 32) / 20; }

int build_tiles(Ctx& c) {
    // tile shape of the current variant (k extent x j rows)
    if (int rc = make_tiles(c, qsrk::tile_k(c.variant), qsrk::tile_j(c.variant), c.ch, &c.d_tiles, &c.tiles_cap,
                            &c.ntiles, &c.tile_voxels, &c.h_tiles))
        return rc;
    if (!fusable(c)) {
        c.ntiles2 = 0;
        c.tile_voxels2 = 0;
        return QSR_OK;
    }
    return make_tiles(c, qsrk::X2_OUT_K, qsrk::X2_OUT_J, c.ch2, &c.d_tiles2, &c.tiles2_cap, &c.ntiles2,
                      &c.tile_voxels2, nullptr);
}

// Box-face voxels that can change in a step (the face kernel's live set: evolving
// and not a frozen strong shell, or a strong shell re-imposed), appended as
// (box, i, j, k) in the face order i-faces, j-faces, k-faces.  Droplets that
// keep clear of their box faces contribute none and the face pass is skipped.
long long append_live_faces(const Box& b, int box, std::vector<int4>& list) {
    const qsr_params& p = b.p;
    const bool skipshell = p.W == 0.0, bc = !(p.W > 0.0);
    const size_t before = list.size();
    auto visit = [&](int i, int j, int k) {
        const int16_t* m = &b.masks.step[(size_t(i) * size_t(p.Ny) + size_t(j)) * 8];
        const bool evolve = k >= m[0] && k <= m[1];
        const bool shell = (k >= m[2] && k <= m[3]) && !(k >= m[4] && k <= m[5]);
        if ((evolve && !(skipshell && shell)) || (bc && shell)) list.push_back(make_int4(box, i, j, k));
    };
    for (int i : {0, p.Nx - 1})
        for (int j = 0; j < p.Ny; ++j)
            for (int k = 0; k < p.Nz; ++k) visit(i, j, k);
    for (int j : {0, p.Ny - 1})
        for (int i = 1; i < p.Nx - 1; ++i)
            for (int k = 0; k < p.Nz; ++k) visit(i, j, k);
    for (int k : {0, p.Nz - 1})
        for (int i = 1; i < p.Nx - 1; ++i)
            for (int j = 1; j < p.Ny - 1; ++j) visit(i, j, k);
    return (long long)(list.size() - before);
}

Ctx* as_ctx(void* p) { return static_cast<Ctx*>(p); }

int create_impl(const qsr_params* params, int n, int device, void** out) {
    if (!params || n < 1 || !out) return fail(QSR_E_ARG, "null argument or n_boxes < 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(QSR_E_NODEVICE, "no HIP device");
    if (device < 0 || device >= ndev) return fail(QSR_E_ARG, "device index out of range");
    for (int i = 0; i < n; ++i) {
        const int rc = check_params(params[i], i);
        if (rc) return rc;
    }
    auto c = std::make_unique<Ctx>();
    c->device = device;
    HIP_TRY(hipSetDevice(device));
    if (device < kMaxDevices && !g_scratch[device]) {
        void* addr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        g_scratch[device] = static_cast<double*>(addr);
    }
    HIP_TRY(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&c->ev[0]));
    HIP_TRY(hipEventCreate(&c->ev[1]));
    long long off = 0, rows_off = 0;
    std::vector<int4> face_list;
    std::vector<qsrk::RowMask> step_rows, obs_rows;
    c->boxes.resize(n);
    c->active_voxels = 0;
    c->box_voxels = 0;
    for (int i = 0; i < n; ++i) {
        Box& b = c->boxes[i];
        b.p = params[i];
        const qsr_params& p = b.p;
        b.m = p.dt / p.gamma;
        b.g = qsrh::make_geometry(p.Nx, p.Ny, p.Nz, p.dx, p.dy, p.dz, p.radius_index, p.center_index);
        b.nvox = (long long)p.Nx * p.Ny * p.Nz;
        b.off = off;
        off += (b.nvox + 63) / 64 * 64;
        b.rows_off = rows_off;
        rows_off += (long long)p.Nx * p.Ny;
        if (!p.periodic) {
            std::string err;
            if (!qsrh::build_row_masks(b.g, p.W, b.masks, err)) return fail(QSR_E_GEOMETRY, err);
            b.thickness = qsrh::shell_thickness(b.g, b.masks.shell_count);
            b.area = qsrh::surface_area(b.g);
            b.nface = append_live_faces(b, i, face_list);
            c->active_voxels += b.masks.active_count;
        } else {
            b.masks.step.assign(size_t(p.Nx) * p.Ny * 8, 0);
            b.masks.obs.assign(size_t(p.Nx) * p.Ny * 8, 0);
            b.nface = 0;
            c->active_voxels += b.nvox;
        }
        c->box_voxels += b.nvox;
        const size_t rows = size_t(p.Nx) * p.Ny;
        for (size_t r = 0; r < rows; ++r) {
            qsrk::RowMask s{}, o{};
            for (int t = 0; t < 8; ++t) {
                s.v[t] = b.masks.step[r * 8 + t];
                o.v[t] = b.masks.obs[r * 8 + t];
            }
            step_rows.push_back(s);
            obs_rows.push_back(o);
        }
        b.last_T = b.T_at(0);
        b.last_S = b.S_at(b.last_T);
        if (p.L2 != 0.0 || p.L3 != 0.0) c->full = true;
        if (p.L2 != 0.0) c->elastic |= 1;
        if (p.L3 != 0.0) c->elastic |= 2;
        if (p.noise_sigma > 0.0) c->noise = true;
    }
    c->cs = std::max<long long>(off, 64);
    c->nface = (long long)face_list.size();
    for (int s = 0; s < 2; ++s) {
        HIP_TRY(hipMalloc(&c->q[s], sizeof(float) * 5 * c->cs));
        HIP_TRY(hipMemsetAsync(c->q[s], 0, sizeof(float) * 5 * c->cs, c->st));
    }
    c->h_boxes.resize(n);
    for (int i = 0; i < n; ++i) c->h_boxes[i] = make_boxdev(c->boxes[i]);
    HIP_TRY(hipMalloc(&c->d_boxes, sizeof(qsrk::BoxDev) * n));
    HIP_TRY(hipMemcpy(c->d_boxes, c->h_boxes.data(), sizeof(qsrk::BoxDev) * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_step_rows, sizeof(qsrk::RowMask) * step_rows.size()));
    HIP_TRY(hipMalloc(&c->d_obs_rows, sizeof(qsrk::RowMask) * obs_rows.size()));
    HIP_TRY(hipMemcpy(c->d_step_rows, step_rows.data(), sizeof(qsrk::RowMask) * step_rows.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_obs_rows, obs_rows.data(), sizeof(qsrk::RowMask) * obs_rows.size(), hipMemcpyHostToDevice));
    {
        std::vector<long long> vp{0};
        for (const Box& b : c->boxes) vp.push_back(vp.back() + b.nvox);
        c->nvox_total = vp.back();
        HIP_TRY(hipMalloc(&c->d_vox_prefix, sizeof(long long) * vp.size()));
        HIP_TRY(hipMemcpy(c->d_vox_prefix, vp.data(), sizeof(long long) * vp.size(), hipMemcpyHostToDevice));
    }
    if (!face_list.empty()) {
        HIP_TRY(hipMalloc(&c->d_face_list, sizeof(int4) * face_list.size()));
        HIP_TRY(hipMemcpy(c->d_face_list, face_list.data(), sizeof(int4) * face_list.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&c->d_dyn, sizeof(qsrk::StepDyn) * kDynChunk * n));
    HIP_TRY(hipHostMalloc((void**)&c->h_dyn, sizeof(qsrk::StepDyn) * kDynChunk * n, hipHostMallocDefault));
    HIP_TRY(hipMalloc(&c->d_out, sizeof(double) * 32));
    // planes per tile: aim at >= 2048 tiles (8 per CU) on the largest box
    long long maxvox = 0;
    for (const Box& b : c->boxes) maxvox = std::max(maxvox, b.nvox);
    c->ch = 64;
    while (c->ch > 8 && maxvox / ((long long)qsrk::TILE_THREADS * c->ch) * (long long)n < 2048) c->ch /= 2;
    const int rc = build_tiles(*c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->st));
    *out = c.release();
    return QSR_OK;
}

int ensure_partials(Ctx& c, long long doubles) {
    if (doubles <= c.part_cap) return QSR_OK;
    discard(hipFree(c.d_part));
    c.d_part = nullptr;
    HIP_TRY(hipMalloc(&c.d_part, sizeof(double) * doubles));
    c.part_cap = doubles;
    return QSR_OK;
}

// Enqueue steps [0, nsteps) of the current chunk (dyn rows already uploaded).  Pairs of
// steps go to the two-step kernel where the context allows it (one buffer swap per
// pair); even_swaps keeps the number of swaps of the same parity as nsteps single steps
// (an even number of pairs), as a replayed graph needs.
int enqueue_steps(Ctx& c, int nsteps, bool even_swaps = false) {
    const int nb = int(c.boxes.size());
    int s = 0;
    if (fusable(c) && c.ntiles2 > 0) {
        int pairs = nsteps / 2;
        if (even_swaps) pairs &= ~1;
        for (int p = 0; p < pairs; ++p, s += 2) {
            const qsrk::StepDyn* dyn = c.d_dyn + (size_t)s * nb;
            HIP_TRY(qsrk::launch_step2(c.ntiles2, c.st, c.q[c.cur], c.q[1 - c.cur], c.cs, c.d_boxes, c.d_tiles2, dyn,
                                       dyn + nb, c.d_step_rows, c.ch2, c.cur));
            c.cur = 1 - c.cur;
        }
    }
    for (; s < nsteps; ++s) {
        const float* qin = c.q[c.cur];
        float* qout = c.q[1 - c.cur];
        const qsrk::StepDyn* dyn = c.d_dyn + (size_t)s * nb;
        if (c.ntiles > 0)
            HIP_TRY(qsrk::launch_step(c.elastic, c.variant, c.noise, c.ntiles, c.st, qin, qout, c.cs, c.d_boxes, c.d_tiles, dyn, nb,
                                      c.d_step_rows, c.ch, nullptr, c.cur));
        HIP_TRY(qsrk::launch_faces(c.nface, c.st, qin, qout, c.cs, c.d_boxes, dyn, c.d_face_list, c.d_step_rows));
        c.cur = 1 - c.cur;
    }
    return QSR_OK;
}

int fill_dyn(Ctx& c, int nsteps) {
    const int nb = int(c.boxes.size());
    for (int s = 0; s < nsteps; ++s)
        for (int b = 0; b < nb; ++b) c.h_dyn[(size_t)s * nb + b] = c.boxes[b].dyn_at(c.boxes[b].iter + s);
    HIP_TRY(hipMemcpyAsync(c.d_dyn, c.h_dyn, sizeof(qsrk::StepDyn) * nsteps * nb, hipMemcpyHostToDevice, c.st));
    return QSR_OK;
}

void advance_iters(Ctx& c, int nsteps) {
    for (Box& b : c.boxes) {
        if (!b.active) continue;
        b.iter += nsteps;
        b.last_T = b.T_at(b.iter - 1);
        b.last_S = b.S_at(b.last_T);
    }
}

int check_uploaded(Ctx& c) {
    for (size_t i = 0; i < c.boxes.size(); ++i)
        if (!c.boxes[i].uploaded) return fail(QSR_E_STATE, "box " + std::to_string(i) + " has no field uploaded");
    return QSR_OK;
}

}  // namespace

extern "C" {

const char* qsr_last_error(void) { return g_err.c_str(); }

int qsr_device_count(int* n) {
    if (!n) return fail(QSR_E_ARG, "null n");
    if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
    return QSR_OK;
}

int qsr_create(const qsr_params* p, int device, void** ctx) { return create_impl(p, 1, device, ctx); }
int qsr_create_batch(const qsr_params* p, int n_boxes, int device, void** ctx) {
    return create_impl(p, n_boxes, device, ctx);
}

int qsr_destroy(void* ctx) {
    if (!ctx) return QSR_OK;
    Ctx* c = as_ctx(ctx);
    discard(hipSetDevice(c->device));
    discard(hipStreamSynchronize(c->st));
    delete c;
    return QSR_OK;
}

int qsr_upload_q(void* ctx, int box, const double* aos5, size_t n_voxels) {
    Ctx* c = as_ctx(ctx);
    if (!c || !aos5 || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    Box& b = c->boxes[box];
    if ((long long)n_voxels != b.nvox) return fail(QSR_E_ARG, "n_voxels does not match Nx*Ny*Nz");
    HIP_TRY(hipSetDevice(c->device));
    double* stage = nullptr;
    HIP_TRY(hipMalloc(&stage, sizeof(double) * 5 * n_voxels));
    hipError_t e = hipMemcpyAsync(stage, aos5, sizeof(double) * 5 * n_voxels, hipMemcpyHostToDevice, c->st);
    if (e == hipSuccess) e = qsrk::launch_aos_to_soa(c->st, stage, c->q[0], c->q[1], c->cs, b.off, b.nvox);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    discard(hipFree(stage));
    if (e != hipSuccess) return fail(QSR_E_HIP, std::string("upload: ") + hipGetErrorString(e));
    b.uploaded = true;
    return QSR_OK;
}

int qsr_download_q(void* ctx, int box, double* aos5, size_t n_voxels) {
    Ctx* c = as_ctx(ctx);
    if (!c || !aos5 || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    Box& b = c->boxes[box];
    if ((long long)n_voxels != b.nvox) return fail(QSR_E_ARG, "n_voxels does not match Nx*Ny*Nz");
    HIP_TRY(hipSetDevice(c->device));
    double* stage = nullptr;
    HIP_TRY(hipMalloc(&stage, sizeof(double) * 5 * n_voxels));
    hipError_t e = qsrk::launch_soa_to_aos(c->st, c->q[c->cur], stage, c->cs, b.off, b.nvox);
    if (e == hipSuccess) e = hipMemcpyAsync(aos5, stage, sizeof(double) * 5 * n_voxels, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    discard(hipFree(stage));
    if (e != hipSuccess) return fail(QSR_E_HIP, std::string("download: ") + hipGetErrorString(e));
    return QSR_OK;
}

int qsr_download_nematic(void* ctx, int box, double* out4, size_t n_voxels) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out4 || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    Box& b = c->boxes[box];
    if ((long long)n_voxels != b.nvox) return fail(QSR_E_ARG, "n_voxels does not match Nx*Ny*Nz");
    HIP_TRY(hipSetDevice(c->device));
    double* stage = nullptr;
    HIP_TRY(hipMalloc(&stage, sizeof(double) * 4 * n_voxels));
    hipError_t e = qsro::launch_nematic_aos(c->st, c->q[c->cur], c->cs, b.off, b.nvox, stage);
    if (e == hipSuccess) e = hipMemcpyAsync(out4, stage, sizeof(double) * 4 * n_voxels, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    discard(hipFree(stage));
    if (e != hipSuccess) return fail(QSR_E_HIP, std::string("nematic: ") + hipGetErrorString(e));
    return QSR_OK;
}

int qsr_step(void* ctx, int nsteps) {
    Ctx* c = as_ctx(ctx);
    if (!c || nsteps < 0) return fail(QSR_E_ARG, "bad argument");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (nsteps == 0) return QSR_OK;
    HIP_TRY(hipEventRecord(c->ev[0], c->st));
    int done = 0;
    while (done < nsteps) {
        const int n = std::min(kDynChunk, nsteps - done);
        HIP_TRY(hipStreamSynchronize(c->st));  // pinned staging is reused per chunk
        if (int rc = fill_dyn(*c, n)) return rc;
        if (int rc = enqueue_steps(*c, n)) return rc;
        advance_iters(*c, n);
        done += n;
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->st));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_ms = double(ms) / nsteps;
    return QSR_OK;
}

int qsr_step_graph(void* ctx, int nsteps, int chunk) {
    Ctx* c = as_ctx(ctx);
    if (!c || nsteps < 0 || chunk < 2 || (chunk & 1) || chunk > kDynChunk)
        return fail(QSR_E_ARG, "chunk must be even, 2..512");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (nsteps == 0) return QSR_OK;
    if (!c->gexec || c->graph_chunk != chunk || c->graph_cur != c->cur) {
        if (c->gexec) discard(hipGraphExecDestroy(c->gexec));
        if (c->graph) discard(hipGraphDestroy(c->graph));
        c->gexec = nullptr;
        c->graph = nullptr;
        HIP_TRY(hipStreamSynchronize(c->st));
        const int save = c->cur;
        HIP_TRY(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
        const int rc = enqueue_steps(*c, chunk, true);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(c->st, &g);
        c->cur = save;
        if (rc) return rc;
        if (e != hipSuccess) return fail(QSR_E_HIP, std::string("capture: ") + hipGetErrorString(e));
        c->graph = g;
        HIP_TRY(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
        c->graph_chunk = chunk;
        c->graph_cur = c->cur;
    }
    HIP_TRY(hipEventRecord(c->ev[0], c->st));
    int done = 0;
    while (nsteps - done >= chunk) {
        HIP_TRY(hipStreamSynchronize(c->st));
        if (int rc = fill_dyn(*c, chunk)) return rc;
        HIP_TRY(hipGraphLaunch(c->gexec, c->st));
        advance_iters(*c, chunk);  // chunk is even: buffer roles are unchanged
        done += chunk;
    }
    if (done < nsteps) {
        HIP_TRY(hipStreamSynchronize(c->st));
        const int n = nsteps - done;
        if (int rc = fill_dyn(*c, n)) return rc;
        if (int rc = enqueue_steps(*c, n)) return rc;
        advance_iters(*c, n);
        // an odd remainder flips the buffers; the graph is re-captured next call
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->st));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_ms = double(ms) / nsteps;
    return QSR_OK;
}

int qsr_profile_steps(void* ctx, int nsteps, double* main_ms, double* face_ms) {
    Ctx* c = as_ctx(ctx);
    if (!c || nsteps < 1 || nsteps > kDynChunk || !main_ms || !face_ms) return fail(QSR_E_ARG, "bad argument");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (int rc = fill_dyn(*c, nsteps)) return rc;
    std::vector<hipEvent_t> ev(3 * size_t(nsteps));
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    const int nb = int(c->boxes.size());
    // a context that pairs its steps is timed per two-step launch (nsteps / 2 of them)
    const bool pairs = fusable(*c) && c->ntiles2 > 0 && nsteps >= 2;
    const int nl = pairs ? nsteps / 2 : nsteps;
    for (int l = 0; l < nl; ++l) {
        const float* qin = c->q[c->cur];
        float* qout = c->q[1 - c->cur];
        const qsrk::StepDyn* dyn = c->d_dyn + (size_t)(pairs ? 2 * l : l) * nb;
        HIP_TRY(hipEventRecord(ev[3 * l], c->st));
        if (pairs)
            HIP_TRY(qsrk::launch_step2(c->ntiles2, c->st, qin, qout, c->cs, c->d_boxes, c->d_tiles2, dyn, dyn + nb,
                                       c->d_step_rows, c->ch2, c->cur));
        else if (c->ntiles > 0)
            HIP_TRY(qsrk::launch_step(c->elastic, c->variant, c->noise, c->ntiles, c->st, qin, qout, c->cs, c->d_boxes, c->d_tiles,
                                      dyn, nb, c->d_step_rows, c->ch, nullptr, c->cur));
        HIP_TRY(hipEventRecord(ev[3 * l + 1], c->st));
        HIP_TRY(qsrk::launch_faces(c->nface, c->st, qin, qout, c->cs, c->d_boxes, dyn, c->d_face_list,
                                   c->d_step_rows));
        HIP_TRY(hipEventRecord(ev[3 * l + 2], c->st));
        c->cur = 1 - c->cur;
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    advance_iters(*c, pairs ? 2 * nl : nl);
    double tm = 0, tf = 0;
    for (int l = 0; l < nl; ++l) {
        float a = 0, b = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[3 * l], ev[3 * l + 1]));
        HIP_TRY(hipEventElapsedTime(&b, ev[3 * l + 1], ev[3 * l + 2]));
        tm += a;
        tf += b;
    }
    for (auto& e : ev) discard(hipEventDestroy(e));
    *main_ms = tm / nl;
    *face_ms = tf / nl;
    return QSR_OK;
}

int qsr_autotune(void* ctx, int trial_steps, int* best_variant, int* best_ch, double* best_ms) {
    Ctx* c = as_ctx(ctx);
    if (!c || trial_steps < 1) return fail(QSR_E_ARG, "bad argument");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    if (int rc = fill_dyn(*c, 1)) return rc;
    const int nb = int(c->boxes.size());
    int maxNx = 0;
    for (const Box& b : c->boxes) maxNx = std::max(maxNx, b.p.Nx);
    const int nvar = 2;  // tile shapes 64 x 4 and 32 x 8
    // planes per tile: the tile count against the 4 x 256 resident blocks matters as much as the
    // lead-in planes per tile (at 256^3, R = 120: 18 planes give 2,028 tiles, just under two
    // rounds, and run 3 % faster than 17 planes' 2,140), so the candidates are dense up to 24
    std::vector<int> chs;
    for (int ch : {8, 10, 12, 14, 16, 18, 20, 24, 32, 64}) {
        if (ch > 8 && ch / 2 >= maxNx) break;
        chs.push_back(ch);
    }
    // Three sweeps over all candidates, each candidate scored by its fastest sweep: a single
    // short sweep lets clock ramp-up and transients outweigh the few per cent between choices.
    constexpr int kSweeps = 3;
    std::vector<double> per_min(chs.size() * nvar, 1e300);
    for (int sweep = 0; sweep < kSweeps; ++sweep) {
        for (size_t ic = 0; ic < chs.size(); ++ic) {
            c->ch = chs[ic];
            for (int v = 0; v < nvar; ++v) {
                c->variant = v;
                if (int rc = build_tiles(*c)) return rc;  // the tile table depends on the shape
                // trial launches read the current buffer and write the spare one: the state does not advance
                for (int s = -1; s < trial_steps; ++s) {
                    if (s == 0) HIP_TRY(hipEventRecord(c->ev[0], c->st));
                    if (c->ntiles > 0)
                        HIP_TRY(qsrk::launch_step(c->elastic, c->variant, c->noise, c->ntiles, c->st, c->q[c->cur],
                                                  c->q[1 - c->cur], c->cs, c->d_boxes, c->d_tiles, c->d_dyn, nb,
                                                  c->d_step_rows, c->ch, nullptr));
                }
                HIP_TRY(hipEventRecord(c->ev[1], c->st));
                HIP_TRY(hipEventSynchronize(c->ev[1]));
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                double& m = per_min[ic * nvar + size_t(v)];
                m = std::min(m, double(ms) / trial_steps);
            }
        }
    }
    double best = 1e300;
    int bv = c->variant, bch = c->ch;
    for (size_t ic = 0; ic < chs.size(); ++ic)
        for (int v = 0; v < nvar; ++v)
            if (per_min[ic * nvar + size_t(v)] < best) {
                best = per_min[ic * nvar + size_t(v)];
                bv = v;
                bch = chs[ic];
            }
    c->variant = bv;
    c->ch = bch;
    if (int rc = build_tiles(*c)) return rc;
    if (fusable(*c)) {
        // planes per tile of the two-step kernel, scored the same way (per step: launch / 2)
        if (int rc = fill_dyn(*c, 2)) return rc;
        std::vector<double> m2(chs.size(), 1e300);
        for (int sweep = 0; sweep < kSweeps; ++sweep)
            for (size_t ic = 0; ic < chs.size(); ++ic) {
                c->ch2 = chs[ic];
                if (int rc = build_tiles(*c)) return rc;
                for (int s = -1; s < trial_steps; ++s) {
                    if (s == 0) HIP_TRY(hipEventRecord(c->ev[0], c->st));
                    HIP_TRY(qsrk::launch_step2(c->ntiles2, c->st, c->q[c->cur], c->q[1 - c->cur], c->cs, c->d_boxes,
                                               c->d_tiles2, c->d_dyn, c->d_dyn + nb, c->d_step_rows, c->ch2, c->cur));
                }
                HIP_TRY(hipEventRecord(c->ev[1], c->st));
                HIP_TRY(hipEventSynchronize(c->ev[1]));
                float ms = 0.f;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                m2[ic] = std::min(m2[ic], double(ms) / (2.0 * trial_steps));
            }
        size_t bi = 0;
        for (size_t ic = 1; ic < chs.size(); ++ic)
            if (m2[ic] < m2[bi]) bi = ic;
        c->ch2 = chs[bi];
        if (int rc = build_tiles(*c)) return rc;
        best = std::min(best, m2[bi]);
    }
    if (c->gexec) discard(hipGraphExecDestroy(c->gexec));
    if (c->graph) discard(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    if (best_variant) *best_variant = bv;
    if (best_ch) *best_ch = bch;
    if (best_ms) *best_ms = best;
    return QSR_OK;
}

int qsr_set_box_active(void* ctx, int box, int active) { return qsr_set_boxes_active(ctx, &box, 1, active); }

int qsr_set_boxes_active(void* ctx, const int* boxes, int n, int active) {
    Ctx* c = as_ctx(ctx);
    if (!c || !boxes || n < 0) return fail(QSR_E_ARG, "bad argument");
    for (int i = 0; i < n; ++i)
        if (boxes[i] < 0 || boxes[i] >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    const bool on = active != 0;
    bool changed = false;
    for (int i = 0; i < n; ++i) changed |= c->boxes[size_t(boxes[i])].active != on;
    if (!changed) return QSR_OK;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; ++i) {
        Box& b = c->boxes[size_t(boxes[i])];
        if (b.active == on) continue;
        if (!on) {
            // both ping-pong buffers hold the held state, whichever one later steps leave current
            for (int comp = 0; comp < 5; ++comp) {
                const size_t o = size_t(comp) * size_t(c->cs) + size_t(b.off);
                HIP_TRY(hipMemcpyAsync(c->q[1 - c->cur] + o, c->q[c->cur] + o, sizeof(float) * size_t(b.nvox),
                                       hipMemcpyDeviceToDevice, c->st));
            }
        }
        b.active = on;
        qsrk::BoxDev& d = c->h_boxes[size_t(boxes[i])];
        d.flags = on ? (d.flags & ~qsrk::F_INACTIVE) : (d.flags | qsrk::F_INACTIVE);
        HIP_TRY(hipMemcpyAsync(c->d_boxes + boxes[i], &d, sizeof(qsrk::BoxDev), hipMemcpyHostToDevice, c->st));
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    if (int rc = build_tiles(*c)) return rc;
    if (c->gexec) discard(hipGraphExecDestroy(c->gexec));
    if (c->graph) discard(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    return QSR_OK;
}

int qsr_set_step_fusion(void* ctx, int on, int planes_per_tile) {
    Ctx* c = as_ctx(ctx);
    if (!c || planes_per_tile > 4096) return fail(QSR_E_ARG, "planes_per_tile must be <= 4096");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    c->fuse = on != 0;
    if (planes_per_tile > 0) c->ch2 = planes_per_tile;
    if (int rc = build_tiles(*c)) return rc;
    if (c->gexec) discard(hipGraphExecDestroy(c->gexec));
    if (c->graph) discard(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    return QSR_OK;
}

int qsr_step_fusion_info(void* ctx, int* active, int* planes_per_tile, uint64_t* tile_voxels) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad argument");
    const bool on = fusable(*c) && c->ntiles2 > 0;
    if (active) *active = on ? 1 : 0;
    if (planes_per_tile) *planes_per_tile = on ? c->ch2 : 0;
    if (tile_voxels) *tile_voxels = on ? uint64_t(c->tile_voxels2) : 0;
    return QSR_OK;
}

int qsr_set_tuning(void* ctx, int variant, int planes_per_tile) {
    Ctx* c = as_ctx(ctx);
    if (!c || variant < 0 || variant > 1 || planes_per_tile < 1 || planes_per_tile > 4096)
        return fail(QSR_E_ARG, "variant must be 0/1 and planes_per_tile in [1, 4096]");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->st));
    c->variant = variant;
    c->ch = planes_per_tile;
    if (int rc = build_tiles(*c)) return rc;
    if (c->gexec) discard(hipGraphExecDestroy(c->gexec));
    if (c->graph) discard(hipGraphDestroy(c->graph));
    c->gexec = nullptr;
    c->graph = nullptr;
    return QSR_OK;
}

int qsr_step_imex(void* ctx, int nsteps, double L_stab, int jacobi_iters) {
    Ctx* c = as_ctx(ctx);
    if (!c || nsteps < 0 || jacobi_iters < 0 || !std::isfinite(L_stab)) return fail(QSR_E_ARG, "bad argument");
    // the reference falls back to the explicit update without a stabiliser (QSR.cu:2721, 2750-2751)
    if (!(L_stab > 0.0) || jacobi_iters == 0) return qsr_step(ctx, nsteps);
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (nsteps == 0) return QSR_OK;
    const int nb = int(c->boxes.size());
    if (!c->q_rhs) HIP_TRY(hipMalloc(&c->q_rhs, sizeof(float) * 5 * c->cs));  // also made by euler_residual
    if (!c->d_imex) {
        HIP_TRY(hipMalloc(&c->d_imex, sizeof(float4) * nb));
        HIP_TRY(hipMalloc(&c->d_jcoef, sizeof(float4) * nb));
    }
    std::vector<float4> im(nb), jc(nb);
    for (int b = 0; b < nb; ++b) {
        const qsr_params& p = c->boxes[b].p;
        const double alpha = (p.dt / p.gamma) * L_stab;  // QSR.cu:766, 2729
        const double ax = alpha / (p.dx * p.dx), ay = alpha / (p.dy * p.dy), az = alpha / (p.dz * p.dz);
        im[b] = make_float4(float(ax), float(ay), float(az), 1.0f);
        jc[b] = make_float4(float(ax), float(ay), float(az), float(1.0 / (1.0 + 2.0 * (ax + ay + az))));
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    HIP_TRY(hipMemcpy(c->d_imex, im.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_jcoef, jc.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
    if (jacobi_iters != c->jflag_n) {
        // the flag rings start cleared; each sweep keeps the next step's row cleared after that
        discard(hipFree(c->d_jflags));
        c->d_jflags = nullptr;
        HIP_TRY(hipMalloc(&c->d_jflags, sizeof(unsigned) * 2 * jacobi_iters));
        HIP_TRY(hipMemsetAsync(c->d_jflags, 0, sizeof(unsigned) * 2 * jacobi_iters, c->st));
        c->jflag_n = jacobi_iters;
        c->jflag_row = 0;
    }
    bool any_bc = false;  // strong anchoring re-imposed after the sweeps (QSR.cu:2753)
    for (const qsrk::BoxDev& d : c->h_boxes) any_bc |= (d.flags & qsrk::F_BC) != 0;
    HIP_TRY(hipEventRecord(c->ev[0], c->st));
    int done = 0;
    while (done < nsteps) {
        const int n = std::min(kDynChunk, nsteps - done);
        HIP_TRY(hipStreamSynchronize(c->st));
        if (int rc = fill_dyn(*c, n)) return rc;
        for (int s = 0; s < n; ++s) {
            const qsrk::StepDyn* dyn = c->d_dyn + (size_t)s * nb;
            if (c->ntiles > 0)
                HIP_TRY(qsrk::launch_step(c->elastic, c->variant, false, c->ntiles, c->st, c->q[c->cur], c->q_rhs, c->cs,
                                          c->d_boxes, c->d_tiles, dyn, nb, c->d_step_rows, c->ch, c->d_imex));
            int src = c->cur;
            HIP_TRY(qsrk::launch_face_copy(c->nface, c->st, c->q[src], c->q[1 - src], c->cs, c->d_boxes,
                                           c->d_face_list));
            for (int it = 0; it < jacobi_iters; ++it) {
                if (c->ntiles > 0)
                    HIP_TRY(qsrk::launch_jacobi(c->variant, c->ntiles, c->st, c->q_rhs, c->q[src], c->q[1 - src], c->cs,
                                                c->d_boxes, c->d_tiles, c->d_step_rows, c->ch, c->d_jcoef,
                                                c->d_jflags, it, jacobi_iters, c->jflag_row));
                src = 1 - src;
            }
            c->jflag_row ^= 1;
            c->cur = src;
            if (any_bc)
                HIP_TRY(qsrk::launch_bc(c->nvox_total, c->st, c->q[c->cur], c->cs, c->d_boxes, dyn, c->d_vox_prefix,
                                        nb, c->d_step_rows));
        }
        advance_iters(*c, n);
        done += n;
    }
    HIP_TRY(hipEventRecord(c->ev[1], c->st));
    HIP_TRY(hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->last_ms = double(ms) / nsteps;
    return QSR_OK;
}

int qsr_last_step_ms(void* ctx, double* ms) {
    Ctx* c = as_ctx(ctx);
    if (!c || !ms) return fail(QSR_E_ARG, "bad argument");
    *ms = c->last_ms;
    return QSR_OK;
}

int qsr_step_work(void* ctx, uint64_t* active, uint64_t* tile, uint64_t* box) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad argument");
    if (active) *active = uint64_t(c->active_voxels);
    if (tile) *tile = uint64_t(c->tile_voxels);
    if (box) *box = uint64_t(c->box_voxels);
    return QSR_OK;
}

int qsr_set_iteration(void* ctx, int box, int64_t iteration) {
    Ctx* c = as_ctx(ctx);
    if (!c || box < 0 || box >= int(c->boxes.size()) || iteration < 0) return fail(QSR_E_ARG, "bad argument");
    Box& b = c->boxes[box];
    b.iter = iteration;
    b.last_T = b.T_at(iteration > 0 ? iteration - 1 : 0);
    b.last_S = b.S_at(b.last_T);
    return QSR_OK;
}

int qsr_box_info_get(void* ctx, int box, qsr_box_info* out) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    const Box& b = c->boxes[box];
    out->voxels = uint64_t(b.nvox);
    out->shell_voxels = b.masks.shell_count;
    out->active_voxels = b.p.periodic ? uint64_t(b.nvox) : b.masks.active_count;
    out->shell_volume = double(b.masks.shell_count) * b.p.dx * b.p.dy * b.p.dz;
    out->surface_area = b.area;
    out->nominal_thickness = 2.0 * b.g.shell_half;
    out->effective_thickness = b.thickness;
    out->S_eq = qsrh::s_eq(qsrh::bulk_abc(b.p.a, b.p.b, b.p.c, b.p.T_high, b.p.T_star, b.p.modechoice));
    out->iteration = b.iter;
    return QSR_OK;
}

}  // extern "C"

namespace {

// Volume-pass parameters of one box at its last stepped temperature.
qsro::ObsParams obs_params(const Ctx& c, int box, bool skip_order) {
    const Box& b = c.boxes[box];
    const qsr_params& p = b.p;
    qsro::ObsParams P{};
    P.Nx = p.Nx; P.Ny = p.Ny; P.Nz = p.Nz; P.periodic = p.periodic;
    P.off = b.off; P.rows_off = b.rows_off;
    P.dx = p.dx; P.dy = p.dy; P.dz = p.dz;
    P.cx = b.g.cx; P.cy = b.g.cy; P.cz = b.g.cz; P.rx = b.g.rx; P.ry = b.g.ry; P.rz = b.g.rz;
    P.min_spacing = b.g.min_spacing;
    const double dT = b.last_T - p.T_star;
    if (p.modechoice == 1) { P.c2 = (3.0 * p.a * dT) / 2.0; P.c3 = (3.0 * p.b) / 2.0; P.c4 = (9.0 * p.c) / 8.0; }
    else { P.c2 = (3.0 * p.a * dT) / 2.0; P.c3 = (9.0 * p.b) / 2.0; P.c4 = (9.0 * p.c) / 4.0; }
    P.no_L = (p.L1 == 0.0 && p.L2 == 0.0 && p.L3 == 0.0);
    P.kappa = p.periodic && !P.no_L ? p.L1 : p.kappa;
    P.L1 = p.L1; P.L2 = p.L2; P.L3 = p.L3;
    P.W_over_delta = (p.W != 0.0 && b.thickness > 0.0) ? p.W / b.thickness : 0.0;
    P.S_shell = b.last_S;
    P.skip_order = skip_order ? 1 : 0;
    return P;
}

int ensure_many(Ctx& c, long long part_doubles, long long out_doubles) {
    if (part_doubles > c.mpart_cap) {
        discard(hipFree(c.d_mpart));
        c.d_mpart = nullptr;
        c.mpart_cap = 0;
        HIP_TRY(hipMalloc(&c.d_mpart, sizeof(double) * part_doubles));
        c.mpart_cap = part_doubles;
    }
    if (out_doubles > c.mout_cap) {
        discard(hipFree(c.d_mout));
        c.d_mout = nullptr;
        c.mout_cap = 0;
        HIP_TRY(hipMalloc(&c.d_mout, sizeof(double) * out_doubles));
        c.mout_cap = out_doubles;
    }
    return QSR_OK;
}

// Room for the tables of a many-box reduction over n boxes.
int ensure_mtab(Ctx& c, int n) {
    if (n <= c.mtab_cap) return QSR_OK;
    for (void* p : {(void*)c.d_mparams, (void*)c.d_mprefix, (void*)c.d_mlist}) discard(hipFree(p));
    c.d_mparams = nullptr;
    c.d_mprefix = nullptr;
    c.d_mlist = nullptr;
    c.mtab_cap = 0;
    HIP_TRY(hipMalloc(&c.d_mparams, sizeof(qsro::ObsParams) * n));
    HIP_TRY(hipMalloc(&c.d_mprefix, sizeof(long long) * (n + 1)));
    HIP_TRY(hipMalloc(&c.d_mlist, sizeof(int) * n));
    c.mtab_cap = n;
    return QSR_OK;
}

}  // namespace

extern "C" {

int qsr_energies(void* ctx, const int* boxes, int n, double* totals) {
    Ctx* c = as_ctx(ctx);
    if (!c || !boxes || !totals || n < 0) return fail(QSR_E_ARG, "bad argument");
    for (int i = 0; i < n; ++i)
        if (boxes[i] < 0 || boxes[i] >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad box");
    if (n == 0) return QSR_OK;
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    // the droplet boxes in one volume launch and one finalize launch (box m owns the same blocks and
    // partial order as a single-box pass, so the sums are the same bit for bit)
    std::vector<int> periodic, drop;
    std::vector<qsro::ObsParams> P;
    std::vector<long long> prefix{0};
    for (int i = 0; i < n; ++i) {
        const Box& b = c->boxes[boxes[i]];
        if (b.p.periodic) {  // the periodic pass also fills a nematic field: one box at a time
            periodic.push_back(i);
            continue;
        }
        drop.push_back(i);
        P.push_back(obs_params(*c, boxes[i], true));
        prefix.push_back(prefix.back() + qsro::obs_volume_blocks(b.nvox));
    }
    const int nd = int(drop.size());
    std::vector<double> v(size_t(nd) * qsro::NPART, 0.0);
    if (nd > 0) {
        if (int rc = ensure_many(*c, prefix.back() * qsro::NPART, (long long)nd * qsro::NPART)) return rc;
        if (int rc = ensure_mtab(*c, nd)) return rc;
        HIP_TRY(hipMemcpyAsync(c->d_mparams, P.data(), sizeof(qsro::ObsParams) * nd, hipMemcpyHostToDevice, c->st));
        HIP_TRY(hipMemcpyAsync(c->d_mprefix, prefix.data(), sizeof(long long) * (nd + 1), hipMemcpyHostToDevice,
                               c->st));
        HIP_TRY(qsro::launch_obs_droplet_multi(c->st, c->q[c->cur], c->cs, c->d_mparams, c->d_mprefix, nd,
                                               prefix.back(), c->d_step_rows, c->d_obs_rows, c->d_mpart, c->d_mout));
        HIP_TRY(hipMemcpyAsync(v.data(), c->d_mout, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
    }
    for (int m = 0; m < nd; ++m) {
        const double* w = v.data() + (long long)m * qsro::NPART;
        totals[drop[m]] = w[0] + w[1] + w[2];
    }
    const qsr_obs_opts oo = {-1, 0.1, 0.25, -1, -1.0, 0, 0, 1};
    for (int i : periodic) {
        qsr_obs o{};
        if (int rc = qsr_observables(ctx, boxes[i], &oo, &o)) return rc;
        totals[i] = o.total;
    }
    return QSR_OK;
}

int qsr_energies_noalloc(void* ctx, const int* boxes, int n, double* totals) {
    Ctx* c = as_ctx(ctx);
    if (!c || !boxes || !totals || n < 0) return fail(QSR_E_ARG, "bad argument");
    if (n > kNoallocMaxBoxes)
        return fail(QSR_E_ARG, "qsr_energies_noalloc: at most " + std::to_string(kNoallocMaxBoxes) + " boxes per call");
    for (int i = 0; i < n; ++i) {
        if (boxes[i] < 0 || boxes[i] >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad box");
        if (c->boxes[boxes[i]].p.periodic)
            return fail(QSR_E_ARG, "qsr_energies_noalloc: droplet boxes only (box " + std::to_string(boxes[i]) +
                                       " is periodic; use qsr_energies)");
    }
    if (n == 0) return QSR_OK;
    if (int rc = check_uploaded(*c)) return rc;
    if (c->device >= kMaxDevices || !g_scratch[c->device])
        return fail(QSR_E_STATE, "qsr_energies_noalloc: no scratch address for this device");
    HIP_TRY(hipSetDevice(c->device));
    double* part = g_scratch[c->device];
    double* sums = part + kNoallocPart;
    double* tot = sums + (long long)kNoallocMaxBoxes * qsro::NPART;
    const float* q = c->q[c->cur];
    for (int i = 0; i < n; ++i) {
        const Box& b = c->boxes[boxes[i]];
        if (qsro::obs_volume_blocks(b.nvox) * qsro::NPART > kNoallocPart)
            return fail(QSR_E_ARG, "qsr_energies_noalloc: partials exceed the scratch");
        const qsro::ObsParams P = obs_params(*c, boxes[i], true);
        long long nblk = 0;
        // the partial slots are reused box after box: the stream orders each box's pass after the
        // previous box's finalize
        HIP_TRY(qsro::launch_obs_droplet(c->st, q, c->cs, P, c->d_step_rows, c->d_obs_rows, part, &nblk));
        HIP_TRY(qsro::launch_obs_finalize(c->st, part, nblk, qsro::NPART, 1u << 7, sums + (long long)i * qsro::NPART));
    }
    hipLaunchKernelGGL(energy_totals_kernel, dim3((n + 255) / 256), dim3(256), 0, c->st, sums, n, tot);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(totals, tot, sizeof(double) * n, hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    return QSR_OK;
}

int qsr_observables(void* ctx, int box, const qsr_obs_opts* opts, qsr_obs* out) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad argument");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    const Box& b = c->boxes[box];
    const qsr_params& p = b.p;
    qsr_obs_opts o{};
    o.z_slice = -1; o.S_threshold = 0.1; o.charge_cutoff = 0.25; o.xi_z_slice = -1; o.xi_S_threshold = -1;
    o.want_defects = 1; o.want_xi = 1;
    if (opts) o = *opts;
    const int zs = std::min(std::max(o.z_slice < 0 ? p.Nz / 2 : o.z_slice, 0), p.Nz - 1);
    const int xz = std::min(std::max(o.xi_z_slice < 0 ? zs : o.xi_z_slice, 0), p.Nz - 1);
    const double xthr = o.xi_S_threshold < 0 ? o.S_threshold : o.xi_S_threshold;

    qsro::ObsParams P = obs_params(*c, box, o.skip_order != 0);
    const double T = b.last_T;
    const float* q = c->q[c->cur];

    qsr_obs r{};
    r.iteration = b.iter - 1;
    r.T = T;
    r.S_shell = b.last_S;
    r.checker_rel = kNaN;
    r.defect_density = kNaN;
    r.xi_grad_proxy = kNaN;
    r.defect_line_density = kNaN;
    double vol[qsro::NPART];
    long long nblk = (b.nvox + 255) / 256;
    if (int rc = ensure_partials(*c, nblk * 10)) return rc;
    if (!p.periodic) {
        HIP_TRY(qsro::launch_obs_droplet(c->st, q, c->cs, P, c->d_step_rows, c->d_obs_rows, c->d_part, &nblk));
        HIP_TRY(qsro::launch_obs_finalize(c->st, c->d_part, nblk, qsro::NPART, 1u << 7, c->d_out));
        HIP_TRY(hipMemcpyAsync(vol, c->d_out, sizeof(vol), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        r.bulk = vol[0]; r.elastic = vol[1]; r.anchoring = vol[2];
        r.total = r.bulk + r.elastic + r.anchoring;
        r.avg_S = vol[4] > 0 ? vol[3] / vol[4] : 0.0;
        r.max_S = vol[7];
        r.radiality = vol[6] > 0 ? vol[5] / vol[6] : 0.0;
        if (o.skip_order) r.avg_S = r.max_S = r.radiality = kNaN;
        // slice pass: defects, xi and (if max_S >= 0.5) the checkerboard metric
        const bool want_checker = !o.skip_order && r.max_S >= 0.5;
        if (!o.want_defects && !o.want_xi && !want_checker) {
            *out = r;
            return QSR_OK;
        }
        const long long ns = (long long)p.Nx * p.Ny;
        if (ns * 4 > c->nf_slice_cap) {
            discard(hipFree(c->d_nf_slice[0]));
            discard(hipFree(c->d_nf_slice[1]));
            c->d_nf_slice[0] = c->d_nf_slice[1] = nullptr;
            HIP_TRY(hipMalloc(&c->d_nf_slice[0], sizeof(double) * ns * 4));
            HIP_TRY(hipMalloc(&c->d_nf_slice[1], sizeof(double) * ns * 4));
            c->nf_slice_cap = ns * 4;
        }
        HIP_TRY(qsro::launch_obs_slice_nematic(c->st, q, c->cs, P, zs, c->d_nf_slice[0]));
        const double* nfx = c->d_nf_slice[0];
        if (xz != zs) {
            HIP_TRY(qsro::launch_obs_slice_nematic(c->st, q, c->cs, P, xz, c->d_nf_slice[1]));
            nfx = c->d_nf_slice[1];
        }
        long long nb2 = 0;
        HIP_TRY(qsro::launch_obs_slice_stats(c->st, c->d_nf_slice[0], nfx, P, c->d_obs_rows, zs, o.S_threshold,
                                             o.charge_cutoff, xthr, want_checker ? 1 : 0, c->d_part, &nb2));
        HIP_TRY(qsro::launch_obs_finalize(c->st, c->d_part, nb2, qsro::NSLICE, 0u, c->d_out));
        double sl[qsro::NSLICE];
        HIP_TRY(hipMemcpyAsync(sl, c->d_out, sizeof(sl), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        if (o.want_defects) {
            r.defects = uint64_t(sl[0]);
            r.plaquettes = uint64_t(sl[1]);
            r.defect_density = r.plaquettes ? double(r.defects) / double(r.plaquettes) : kNaN;
        }
        if (o.want_xi) {
            r.xi_edges = uint64_t(sl[3]);
            if (r.xi_edges > 0) {
                const double mean_sq = sl[2] / double(r.xi_edges);
                const double eps = 1e-30;
                const double inv = (mean_sq > 0.0) ? 1.0 / std::sqrt(mean_sq + eps) : 1.0 / std::sqrt(eps);
                r.xi_grad_proxy = std::min(inv, 0.5 * double(std::min(p.Nx, p.Ny)));
            }
        }
        if (want_checker && sl[5] > 10 && sl[7] > 10 && sl[9] > 20 && sl[8] != 0.0)
            r.checker_rel = std::abs(sl[4] - sl[6]) / std::abs(sl[8]);
    } else {
        if (b.nvox * 4 > c->nf3_cap) {
            discard(hipFree(c->d_nf3));
            c->d_nf3 = nullptr;
            HIP_TRY(hipMalloc(&c->d_nf3, sizeof(double) * b.nvox * 4));
            c->nf3_cap = b.nvox * 4;
        }
        const qsrh::BulkABC k = qsrh::bulk_abc(p.a, p.b, p.c, T, p.T_star, p.modechoice);
        P.c2 = k.A / 2.0; P.c3 = k.B / 3.0; P.c4 = k.C / 4.0;
        HIP_TRY(qsro::launch_obs_periodic(c->st, q, c->cs, P, c->d_nf3, c->d_part, &nblk));
        HIP_TRY(qsro::launch_obs_finalize(c->st, c->d_part, nblk, qsro::NPART, 1u << 7, c->d_out));
        HIP_TRY(hipMemcpyAsync(vol, c->d_out, sizeof(vol), hipMemcpyDeviceToHost, c->st));
        long long nb3 = 0;
        HIP_TRY(qsro::launch_obs_plaquette3(c->st, c->d_nf3, P, o.S_threshold, o.charge_cutoff, c->d_part, &nb3));
        HIP_TRY(qsro::launch_obs_finalize(c->st, c->d_part, nb3, qsro::NPLAQ, 0u, c->d_out + 16));
        double pl[qsro::NPLAQ];
        HIP_TRY(hipMemcpyAsync(pl, c->d_out + 16, sizeof(pl), hipMemcpyDeviceToHost, c->st));
        HIP_TRY(hipStreamSynchronize(c->st));
        const double V = p.dx * p.dy * p.dz * double(b.nvox);
        r.bulk = vol[0] / V;
        r.elastic = vol[1] / V;
        r.anchoring = 0.0;
        r.total = r.bulk + r.elastic;
        r.radiality = kNaN;
        r.avg_S = vol[4] / double(b.nvox);
        r.max_S = vol[7];
        r.defects = uint64_t(pl[0]);
        r.plaquettes = uint64_t(pl[1]);
        r.defect_density = r.plaquettes ? pl[0] / pl[1] : kNaN;
        r.defect_line_density = V > 0 ? pl[2] / V : kNaN;
        if (vol[3] > 0.0 && vol[2] > 0.0) r.xi_grad_proxy = std::sqrt(vol[2] / vol[3]);
    }
    *out = r;
    return QSR_OK;
}

double qsr_S_eq(double a, double b, double c, double T, double T_star, int modechoice) {
    return qsrh::s_eq(qsrh::bulk_abc(a, b, c, T, T_star, modechoice));
}

int qsr_dt_estimate(double dx, double dy, double dz, double gamma, double a, double b, double c, double T_star,
                    int modechoice, double kappa, double L1, double L2, double L3, double S_ref_used, double S0,
                    double T_a, double T_b, int use_semi_implicit, double L_stab, int jacobi_iters, double out[10]) {
    if (!out) return fail(QSR_E_ARG, "null out");
    qsrh::DtInput in;
    in.dx = dx; in.dy = dy; in.dz = dz; in.gamma = gamma; in.a = a; in.b = b; in.c = c; in.T_star = T_star;
    in.modechoice = modechoice; in.kappa = kappa; in.L1 = L1; in.L2 = L2; in.L3 = L3;
    in.S_ref_used = S_ref_used; in.S0 = S0; in.T_a = T_a; in.T_b = T_b;
    in.semi_implicit = use_semi_implicit != 0; in.L_stab = L_stab; in.jacobi_iters = jacobi_iters;
    const qsrh::DtWindow w = qsrh::stable_dt(in);
    const double v[10] = {w.min_dx, w.S_scale, w.L_eff_total, w.D_total, w.dt_diff_total,
                          w.L_eff_explicit, w.dt_diff, w.bulk_rate, w.dt_bulk, w.dt_max};
    std::memcpy(out, v, sizeof(v));
    return QSR_OK;
}

int qsr_load_qtensor_file(const char* path, int Nx, int Ny, int Nz, double* aos5_out, size_t n_voxels,
                          uint64_t* rows_read) {
    if (!path || !aos5_out || Nx < 1 || Ny < 1 || Nz < 1) return fail(QSR_E_ARG, "bad argument");
    if ((long long)n_voxels != (long long)Nx * Ny * Nz) return fail(QSR_E_ARG, "n_voxels mismatch");
    std::vector<double> Q;
    uint64_t rows = 0;
    std::string err;
    if (!qsrh::load_qtensor_text(path, Nx, Ny, Nz, Q, rows, err)) return fail(QSR_E_ARG, err);
    std::memcpy(aos5_out, Q.data(), sizeof(double) * Q.size());
    if (rows_read) *rows_read = rows;
    return QSR_OK;
}

int qsr_save_qtensor_file(const char* path, const double* aos5, int Nx, int Ny, int Nz, size_t n_voxels) {
    if (!path || !aos5 || Nx < 1 || Ny < 1 || Nz < 1) return fail(QSR_E_ARG, "bad argument");
    if ((long long)n_voxels != (long long)Nx * Ny * Nz) return fail(QSR_E_ARG, "n_voxels mismatch");
    std::string err;
    if (!qsrh::save_qtensor_text(path, aos5, Nx, Ny, Nz, err)) return fail(QSR_E_ARG, err);
    return QSR_OK;
}

int qsr_init_field(const qsr_params* p, int mode, double S0, double noise_amplitude, int64_t random_seed,
                   double* aos5_out, size_t n_voxels) {
    if (!p || !aos5_out) return fail(QSR_E_ARG, "null argument");
    if ((long long)n_voxels != (long long)p->Nx * p->Ny * p->Nz) return fail(QSR_E_ARG, "n_voxels mismatch");
    const qsrh::Geometry g = qsrh::make_geometry(p->Nx, p->Ny, p->Nz, p->dx, p->dy, p->dz, p->radius_index,
                                                 p->center_index);
    std::vector<double> Q;
    if (mode == 2) qsrh::init_isotropic(Q, g, noise_amplitude, random_seed);
    else qsrh::init_director(Q, g, S0, mode == 0 ? 0 : 1, random_seed);
    std::memcpy(aos5_out, Q.data(), sizeof(double) * Q.size());
    return QSR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Internal access for the learned warm-start step (qsr_ctx.hpp)
// ---------------------------------------------------------------------------
namespace qsrc {

int set_error(int code, const std::string& msg) { return fail(code, msg); }

int field_view(void* ctx, int box, FieldView* out) {
    Ctx* c = as_ctx(ctx);
    if (!c || !out || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad context or box");
    FieldView v;
    v.device = c->device;
    v.st = c->st;
    v.cur = c->q[c->cur];
    v.other = c->q[1 - c->cur];
    v.rhs = c->q_rhs;
    v.cs = c->cs;
    v.nvox = c->boxes[box].nvox;
    v.box = c->h_boxes[box];
    v.rows = c->d_step_rows;
    v.host_rows = c->boxes[box].masks.step.empty() ? nullptr : c->boxes[box].masks.step.data();
    v.nboxes = int(c->boxes.size());
    v.uploaded = c->boxes[box].uploaded;
    v.active = c->boxes[box].active;
    *out = v;
    return QSR_OK;
}

int commit_other(void* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad context");
    if (c->boxes.size() != 1) return fail(QSR_E_STATE, "commit_other needs a single-box context");
    c->cur = 1 - c->cur;
    return QSR_OK;
}

int swap_buffers(void* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad context");
    c->cur = 1 - c->cur;
    return QSR_OK;
}

int copy_current_to_other(void* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->q[1 - c->cur], c->q[c->cur], sizeof(float) * 5 * c->cs, hipMemcpyDeviceToDevice,
                           c->st));
    return QSR_OK;
}

int copy_box_other_to_current(void* ctx, int box) {
    Ctx* c = as_ctx(ctx);
    if (!c || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad context or box");
    HIP_TRY(hipSetDevice(c->device));
    const qsrk::BoxDev& B = c->h_boxes[box];
    const long long n = c->boxes[box].nvox;
    for (int comp = 0; comp < 5; ++comp) {
        const long long o = (long long)comp * c->cs + B.off;
        HIP_TRY(hipMemcpyAsync(c->q[c->cur] + o, c->q[1 - c->cur] + o, sizeof(float) * n, hipMemcpyDeviceToDevice,
                               c->st));
    }
    return QSR_OK;
}

int euler_rhs(void* ctx) {
    Ctx* c = as_ctx(ctx);
    if (!c) return fail(QSR_E_ARG, "bad context");
    if (int rc = check_uploaded(*c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    const int nb = int(c->boxes.size());
    if (!c->q_rhs) HIP_TRY(hipMalloc(&c->q_rhs, sizeof(float) * 5 * c->cs));
    if (!c->d_rhs0) {
        std::vector<float4> z(nb, make_float4(0.f, 0.f, 0.f, 1.f));
        HIP_TRY(hipMalloc(&c->d_rhs0, sizeof(float4) * nb));
        HIP_TRY(hipMemcpy(c->d_rhs0, z.data(), sizeof(float4) * nb, hipMemcpyHostToDevice));
    }
    HIP_TRY(hipStreamSynchronize(c->st));
    if (int rc = fill_dyn(*c, 1)) return rc;
    if (c->ntiles > 0)
        HIP_TRY(qsrk::launch_step(c->elastic, c->variant, false, c->ntiles, c->st, c->q[c->cur], c->q_rhs, c->cs,
                                  c->d_boxes, c->d_tiles, c->d_dyn, nb, c->d_step_rows, c->ch, c->d_rhs0));
    return QSR_OK;
}

int box_residual(void* ctx, int box, double* res, long long* count) {
    Ctx* c = as_ctx(ctx);
    if (!c || !res || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad context or box");
    if (!c->q_rhs) return fail(QSR_E_STATE, "no explicit step computed");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = ensure_partials(*c, 4096)) return rc;
    long long nblk = 0;
    HIP_TRY(qsrs::launch_residual(c->st, c->q[c->cur], c->q_rhs, c->cs, c->h_boxes[box], c->d_step_rows, c->d_part,
                                  &nblk));
    HIP_TRY(qsrs::launch_sum_slots(c->st, c->d_part, nblk, 2, c->d_out));
    double r[2] = {0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(r, c->d_out, sizeof(r), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    *res = std::sqrt(r[0]);
    if (count) *count = (long long)r[1];
    return QSR_OK;
}

int box_residuals(void* ctx, const int* boxes, int n, double* res, long long* counts) {
    Ctx* c = as_ctx(ctx);
    if (!c || !boxes || !res || n < 0) return fail(QSR_E_ARG, "bad argument");
    if (!c->q_rhs) return fail(QSR_E_STATE, "no explicit step computed");
    if (n == 0) return QSR_OK;
    HIP_TRY(hipSetDevice(c->device));
    // every box in one launch and one summing launch (box i owns the blocks of a single-box
    // residual pass, so the sums are the same bit for bit)
    std::vector<long long> prefix{0};
    for (int i = 0; i < n; ++i) {
        if (boxes[i] < 0 || boxes[i] >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad box");
        prefix.push_back(prefix.back() + qsrs::residual_blocks(c->boxes[boxes[i]].nvox));
    }
    if (int rc = ensure_many(*c, 2 * prefix.back(), 2LL * n)) return rc;
    if (int rc = ensure_mtab(*c, n)) return rc;
    HIP_TRY(hipMemcpyAsync(c->d_mprefix, prefix.data(), sizeof(long long) * (n + 1), hipMemcpyHostToDevice, c->st));
    HIP_TRY(hipMemcpyAsync(c->d_mlist, boxes, sizeof(int) * n, hipMemcpyHostToDevice, c->st));
    HIP_TRY(qsrs::launch_residual_multi(c->st, c->q[c->cur], c->q_rhs, c->cs, c->d_boxes, c->d_mlist, c->d_mprefix, n,
                                        prefix.back(), c->d_step_rows, c->d_mpart, c->d_mout));
    std::vector<double> v(2 * size_t(n));
    HIP_TRY(hipMemcpyAsync(v.data(), c->d_mout, sizeof(double) * v.size(), hipMemcpyDeviceToHost, c->st));
    HIP_TRY(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; ++i) {
        res[i] = std::sqrt(v[2 * i]);
        if (counts) counts[i] = (long long)v[2 * i + 1];
    }
    return QSR_OK;
}

int euler_residual(void* ctx, int box, double* res, long long* count) {
    Ctx* c = as_ctx(ctx);
    if (!c || !res || box < 0 || box >= int(c->boxes.size())) return fail(QSR_E_ARG, "bad context or box");
    if (int rc = euler_rhs(ctx)) return rc;
    return box_residual(ctx, box, res, count);
}

}  // namespace qsrc
