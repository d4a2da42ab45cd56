// qsr_surrogate.cpp -- the C ABI of the learned warm-start step
// (include/qsr_amd.h, qsr_surr_*): parameters, training on (Q_k, Q_k+N)
// pairs that the plain solver produces, and the accepted-or-rejected
// application to a solver context.  The kernels are in qsr_surr.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/qsr_amd.h"
#include "qsr_ctx.hpp"
#include "qsr_surr.hpp"

namespace {

using namespace qsrs;

int fail(int code, const std::string& msg) { return qsrc::set_error(code, msg); }

#define HIP_TRY(expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) return fail(QSR_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

inline void discard(hipError_t) {}

// splitmix64 (Steele, Lea & Flood 2014): crop sampling and initial weights
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
    int below(int n) { return int(next() % uint64_t(n)); }
};

struct SnapH {
    float* q = nullptr;
    uint8_t* mask = nullptr;
    float* inc = nullptr;  // one explicit step's increment (dataset_add_increment)
    double inc_rms = 0.0;  // its RMS per component over the voxels the residual sums over
    int X = 0, Y = 0, Z = 0;
    long long nvox = 0;
};

void free_snap(SnapH& s) {
    discard(hipFree(s.q));
    discard(hipFree(s.mask));
    discard(hipFree(s.inc));
    s.q = s.inc = nullptr;
    s.mask = nullptr;
}

struct Net {
    int device = 0;
    hipStream_t st = nullptr;
    float *P = nullptr, *G = nullptr, *M = nullptr, *Vm = nullptr;
    Frags fr{};
    void* frag_mem = nullptr;
    long long t = 0;  // Adam steps taken
    bool has_params = false;
    int input_kind = QSR_SURR_INPUT_FIELD;
    double horizon = 1.0;  // N: increment input predicts dQ = N s y
    // training / evaluation workspace
    Work w{};
    long long w_cap = 0;
    void* w_mem = nullptr;
    float* stage = nullptr;  // host batch upload (x5, t5, mask)
    long long stage_cap = 0;
    double* d_out = nullptr;  // 2 doubles per training iteration (loss partial sums)
    long long out_cap = 0;
    // application workspace: the listed boxes' activations packed box after box, and the tile
    // table of their network pass (device, and pinned host memory it is staged in)
    uint16_t *ax = nullptr, *aa1 = nullptr, *aa2 = nullptr;
    long long a_cap = 0;
    int4 *d_tiles = nullptr, *h_tiles = nullptr;
    ConvBox *d_cboxes = nullptr, *h_cboxes = nullptr;
    long long tiles_cap = 0, cboxes_cap = 0;
    hipEvent_t ev_stage = nullptr;  // the last staging copies; waited for before the staging is reused
    bool stage_pending = false;
    bool tiles_all = false;  // list every tile of a box (QSR_SURR_TILES_ALL)
    // dataset
    std::vector<SnapH> snaps;
    Snap* d_snaps = nullptr;  // device copy of the snapshot table
    bool d_snaps_stale = true;  // set whenever the host table changes (add, clear)
    CropJob* d_jobs = nullptr;
    long long jobs_cap = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};

    ~Net() {
        for (SnapH& s : snaps) free_snap(s);
        for (void* p : {(void*)P, (void*)G, (void*)M, (void*)Vm, frag_mem, w_mem, (void*)stage, (void*)d_out,
                        (void*)ax, (void*)aa1, (void*)aa2, (void*)d_snaps, (void*)d_jobs, (void*)d_tiles,
                        (void*)d_cboxes})
            discard(hipFree(p));
        for (void* p : {(void*)h_tiles, (void*)h_cboxes})
            if (p) discard(hipHostFree(p));
        for (hipEvent_t e : ev)
            if (e) discard(hipEventDestroy(e));
        if (ev_stage) discard(hipEventDestroy(ev_stage));
        if (st) discard(hipStreamDestroy(st));
    }
};

Net* as_net(void* p) { return static_cast<Net*>(p); }

int ensure_work(Net& n, long long V) {
    if (V <= n.w_cap) {
        n.w.V = V;
        return QSR_OK;
    }
    discard(hipFree(n.w_mem));
    n.w_mem = nullptr;
    n.w_cap = 0;
    // element sizes per voxel, 256-byte aligned sub-buffers
    const long long b16_8 = V * 8 * 2, b16_32 = V * 32 * 2, f32_8 = V * 8 * 4, f32_32 = V * 32 * 4, f32_1 = V * 4;
    auto al = [](long long b) { return (b + 255) / 256 * 256; };
    const long long sizes[] = {b16_8,  b16_8,  f32_32, f32_32, b16_32, b16_32, b16_32, b16_32, f32_8, f32_8,
                               f32_1,  b16_8,  b16_8,  b16_32, b16_32, b16_32, b16_32, part_floats() * 4,
                               2048 * 2 * 8, f32_8};
    constexpr int NBUF = int(sizeof(sizes) / sizeof(sizes[0]));
    long long total = 0;
    for (long long s : sizes) total += al(s);
    HIP_TRY(hipMalloc(&n.w_mem, total));
    char* p = static_cast<char*>(n.w_mem);
    void* ptrs[NBUF];
    for (int i = 0; i < NBUF; ++i) {
        ptrs[i] = p;
        p += al(sizes[i]);
    }
    Work& w = n.w;
    w.x = (uint16_t*)ptrs[0];
    w.x_cm = (uint16_t*)ptrs[1];
    w.z1 = (float*)ptrs[2];
    w.z2 = (float*)ptrs[3];
    w.a1 = (uint16_t*)ptrs[4];
    w.a1_cm = (uint16_t*)ptrs[5];
    w.a2 = (uint16_t*)ptrs[6];
    w.a2_cm = (uint16_t*)ptrs[7];
    w.pred = (float*)ptrs[8];
    w.tgt = (float*)ptrs[9];
    w.mask = (float*)ptrs[10];
    w.dy = (uint16_t*)ptrs[11];
    w.dy_cm = (uint16_t*)ptrs[12];
    w.dz2 = (uint16_t*)ptrs[13];
    w.dz2_cm = (uint16_t*)ptrs[14];
    w.dz1 = (uint16_t*)ptrs[15];
    w.dz1_cm = (uint16_t*)ptrs[16];
    w.part = (float*)ptrs[17];
    w.red = (double*)ptrs[18];
    w.mod = (float*)ptrs[19];
    w.V = V;
    n.w_cap = V;
    return QSR_OK;
}

int ensure_out(Net& n, long long doubles) {
    if (doubles <= n.out_cap) return QSR_OK;
    discard(hipFree(n.d_out));
    n.d_out = nullptr;
    HIP_TRY(hipMalloc(&n.d_out, sizeof(double) * doubles));
    n.out_cap = doubles;
    return QSR_OK;
}

ConvArgs conv_dims(int nb, int X, int Y, int Z) {
    ConvArgs a{};
    a.n = nb;
    a.X = X;
    a.Y = Y;
    a.Z = Z;
    a.V = (long long)nb * X * Y * Z;
    a.segs_row = (Z + 15) / 16;
    a.nseg = (long long)nb * X * Y * a.segs_row;
    return a;
}

// Forward over the workspace batch; with `train` the pre-activations and the
// channel-major copies the backward pass needs are kept.
int forward(Net& n, int nb, int X, int Y, int Z, bool train) {
    Work& w = n.w;
    ConvArgs a = conv_dims(nb, X, Y, Z);
    a.in = w.x;
    a.frag = n.fr.f1;
    a.bias = n.fr.b1;
    a.z_out = train ? w.z1 : nullptr;
    a.a_out = w.a1;
    a.a_cm = train ? w.a1_cm : nullptr;
    HIP_TRY(launch_conv(8, 2, EPI_HIDDEN, n.st, a));
    a.in = w.a1;
    a.frag = n.fr.f2;
    a.bias = n.fr.b2;
    a.z_out = train ? w.z2 : nullptr;
    a.a_out = w.a2;
    a.a_cm = train ? w.a2_cm : nullptr;
    HIP_TRY(launch_conv(32, 2, EPI_HIDDEN, n.st, a));
    a.in = w.a2;
    a.frag = n.fr.f3;
    a.bias = n.fr.b3;
    a.z_out = nullptr;
    a.a_out = nullptr;
    a.a_cm = nullptr;
    a.pred = w.pred;
    HIP_TRY(launch_conv(32, 1, EPI_PRED, n.st, a));
    return QSR_OK;
}

// Loss partial sums into out[0..1] (sum m (y-t)^2, sum m) and dy.
int loss(Net& n, double* out) {
    long long nblk = 0;
    const float scale = float(2.0 / (5.0 * double(n.w.V)));
    HIP_TRY(launch_loss(n.st, n.w, scale, &nblk));
    HIP_TRY(launch_sum_slots(n.st, n.w.red, nblk, 2, out));
    return QSR_OK;
}

// Backward from dy into the flat gradient G.
int backward(Net& n, int nb, int X, int Y, int Z) {
    Work& w = n.w;
    const long long V = w.V;
    HIP_TRY(launch_wgrad(32, 1, n.st, w.a2_cm, w.dy_cm, X, Y, Z, V, w.part));
    HIP_TRY(launch_wgrad_reduce(32, 1, CH_H, CH_Q, OFF_W3, OFF_B3, n.st, w.part, n.G));
    ConvArgs a = conv_dims(nb, X, Y, Z);
    a.in = w.dy;
    a.frag = n.fr.g3;
    a.z_in = w.z2;
    a.a_out = w.dz2;
    a.a_cm = w.dz2_cm;
    HIP_TRY(launch_conv(8, 2, EPI_BWD, n.st, a));
    HIP_TRY(launch_wgrad(32, 2, n.st, w.a1_cm, w.dz2_cm, X, Y, Z, V, w.part));
    HIP_TRY(launch_wgrad_reduce(32, 2, CH_H, CH_H, OFF_W2, OFF_B2, n.st, w.part, n.G));
    a.in = w.dz2;
    a.frag = n.fr.g2;
    a.z_in = w.z1;
    a.a_out = w.dz1;
    a.a_cm = w.dz1_cm;
    HIP_TRY(launch_conv(32, 2, EPI_BWD, n.st, a));
    HIP_TRY(launch_wgrad(8, 2, n.st, w.x_cm, w.dz1_cm, X, Y, Z, V, w.part));
    HIP_TRY(launch_wgrad_reduce(8, 2, CH_Q, CH_H, OFF_W1, OFF_B1, n.st, w.part, n.G));
    return QSR_OK;
}

int adam(Net& n, double lr) {
    constexpr double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    n.t += 1;
    const double bc1 = 1.0 - std::pow(b1, double(n.t)), bc2 = 1.0 - std::pow(b2, double(n.t));
    HIP_TRY(launch_adam(n.st, n.P, n.G, n.M, n.Vm, NPARAM, float(lr / bc1), float(b1), float(b2), float(eps),
                        float(1.0 / std::sqrt(bc2))));
    HIP_TRY(launch_pack_frags(n.st, n.P, n.fr));
    return QSR_OK;
}

int check_net(Net* n) {
    if (!n) return fail(QSR_E_ARG, "null network");
    if (!n->has_params) return fail(QSR_E_STATE, "network has no parameters (init, set or load them first)");
    return QSR_OK;
}

// One box of a network pass: its view of the context and, for increment and
// step-size input, the RMS s of its increment.
struct Prop {
    qsrc::FieldView v;
    double s_inc;
};

// The tiles of one box for a tile-listed network pass, appended to out as
// (lbox, i0, j0, k0) with i0 a multiple of ch, j0 of TILE_TJ and k0 of TILE_TK:
// every tile holding a voxel within TILE_REACH (in each direction) of a voxel
// inside its row's evolve interval, which holds every voxel a step can change.
// The interval hull of the rows within reach of a tile's rows bounds the tile's
// k range.  Periodic boxes, boxes without host row masks and `all` list every tile.
void box_tiles(const qsrc::FieldView& v, int ch, bool all, int lbox, std::vector<int4>& out) {
    const int X = v.box.Nx, Y = v.box.Ny, Z = v.box.Nz, R = TILE_REACH;
    const bool every = all || (v.box.flags & qsrk::F_PERIODIC) || !v.host_rows;
    const int ntk = (Z + TILE_TK - 1) / TILE_TK;
    for (int i0 = 0; i0 < X; i0 += ch)
        for (int j0 = 0; j0 < Y; j0 += TILE_TJ) {
            int lo = 0, hi = Z - 1;
            if (!every) {
                lo = Z;
                hi = -1;
                const int ia = std::max(0, i0 - R), ib = std::min(X - 1, i0 + ch - 1 + R);
                const int ja = std::max(0, j0 - R), jb = std::min(Y - 1, j0 + TILE_TJ - 1 + R);
                for (int i = ia; i <= ib; ++i)
                    for (int j = ja; j <= jb; ++j) {
                        const int16_t* m = v.host_rows + (size_t(i) * Y + j) * 8;
                        if (m[0] <= m[1]) {
                            lo = std::min(lo, int(m[0]));
                            hi = std::max(hi, int(m[1]));
                        }
                    }
                if (lo > hi) continue;
                lo -= R;
                hi += R;
            }
            for (int kt = 0; kt < ntk; ++kt) {
                const int k0 = kt * TILE_TK;
                if (k0 + TILE_TK - 1 >= lo && k0 <= hi) out.push_back(make_int4(lbox, i0, j0, k0));
            }
        }
}

// Network pass over the listed boxes of one context, on the context's stream: for
// each box the proposal q_old + out_scale * (y, or u (1 + y) under increment and
// gain input) into v.other on the voxels a step can change, from v.cur (and
// v.rhs = E(v.cur)).  The boxes' inputs are packed box after box and each layer is
// one launch over the tiles of every listed box.  With ms_out the pass is timed
// (and waited for).
int propose_list(Net& n, const std::vector<Prop>& props, float out_scale, double* ms_out,
                 long long* voxels_out) {
    if (ms_out) *ms_out = 0.0;
    if (voxels_out) *voxels_out = 0;
    if (props.empty()) return QSR_OK;
    const bool incr = n.input_kind != QSR_SURR_INPUT_FIELD;
    const qsrc::FieldView& v0 = props[0].v;
    const hipStream_t st = v0.st;
    const int nb = int(props.size());
    std::vector<long long> off(static_cast<size_t>(nb));
    long long V = 0;
    for (int b = 0; b < nb; ++b) {
        off[b] = V;
        V += props[b].v.nvox;
    }
    if (V > n.a_cap) {
        for (uint16_t** p : {&n.ax, &n.aa1, &n.aa2}) {
            discard(hipFree(*p));
            *p = nullptr;
        }
        n.a_cap = 0;
        HIP_TRY(hipMalloc(&n.ax, sizeof(uint16_t) * 8 * V));
        HIP_TRY(hipMalloc(&n.aa1, sizeof(uint16_t) * 32 * V));
        HIP_TRY(hipMalloc(&n.aa2, sizeof(uint16_t) * 32 * V));
        n.a_cap = V;
    }
    // planes per tile: the largest of 16, 8 and 4 that lists at least 1024 tiles
    std::vector<int4> tiles;
    int ch = 16;
    for (;;) {
        tiles.clear();
        for (int b = 0; b < nb; ++b) box_tiles(props[b].v, ch, n.tiles_all, b, tiles);
        if (tiles.size() >= 1024 || ch <= 4) break;
        ch /= 2;
    }
    if (tiles.empty()) return QSR_OK;  // no voxel of any listed box can change
    if (voxels_out)
        for (const int4& t : tiles) {
            const qsrk::BoxDev& bx = props[t.x].v.box;
            *voxels_out += (long long)std::min(ch, bx.Nx - t.y) * std::min(TILE_TJ, bx.Ny - t.z) *
                           std::min(TILE_TK, bx.Nz - t.w);
        }
    // stage the tile table and the boxes in pinned memory (once the previous copies are done)
    if (n.stage_pending) {
        HIP_TRY(hipEventSynchronize(n.ev_stage));
        n.stage_pending = false;
    }
    const long long nt = (long long)tiles.size();
    if (nt > n.tiles_cap) {
        discard(hipFree(n.d_tiles));
        if (n.h_tiles) discard(hipHostFree(n.h_tiles));
        n.d_tiles = n.h_tiles = nullptr;
        n.tiles_cap = 0;
        HIP_TRY(hipMalloc(&n.d_tiles, sizeof(int4) * nt));
        HIP_TRY(hipHostMalloc(&n.h_tiles, sizeof(int4) * nt, hipHostMallocDefault));
        n.tiles_cap = nt;
    }
    if (nb > n.cboxes_cap) {
        discard(hipFree(n.d_cboxes));
        if (n.h_cboxes) discard(hipHostFree(n.h_cboxes));
        n.d_cboxes = n.h_cboxes = nullptr;
        n.cboxes_cap = 0;
        HIP_TRY(hipMalloc(&n.d_cboxes, sizeof(ConvBox) * nb));
        HIP_TRY(hipHostMalloc(&n.h_cboxes, sizeof(ConvBox) * nb, hipHostMallocDefault));
        n.cboxes_cap = nb;
    }
    std::copy(tiles.begin(), tiles.end(), n.h_tiles);
    for (int b = 0; b < nb; ++b) {
        const qsrk::BoxDev& bx = props[b].v.box;
        n.h_cboxes[b] = ConvBox{bx.Nx, bx.Ny, bx.Nz, 0, off[b], bx};
    }
    HIP_TRY(hipMemcpyAsync(n.d_tiles, n.h_tiles, sizeof(int4) * nt, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(n.d_cboxes, n.h_cboxes, sizeof(ConvBox) * nb, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(n.ev_stage, st));
    n.stage_pending = true;

    if (ms_out) HIP_TRY(hipEventRecord(n.ev[0], st));
    for (int b = 0; b < nb; ++b) {
        const qsrc::FieldView& v = props[b].v;
        uint16_t* x = n.ax + off[b] * 8;
        if (n.input_kind == QSR_SURR_INPUT_INCREMENT || n.input_kind == QSR_SURR_INPUT_STEP)
            HIP_TRY(launch_pack_increment(st, v.cur, v.rhs, v.cs, v.box, v.rows, float(1.0 / props[b].s_inc), x));
        else
            HIP_TRY(launch_pack_field(st, v.cur, v.cs, v.box.off, v.nvox, x));
    }
    ConvArgs a{};
    a.n = 1;
    a.tiles = n.d_tiles;
    a.boxes = n.d_cboxes;
    a.ntiles = int(nt);
    a.tile_ch = ch;
    a.in = n.ax;
    a.frag = n.fr.f1;
    a.bias = n.fr.b1;
    a.a_out = n.aa1;
    HIP_TRY(launch_conv(8, 2, EPI_HIDDEN, st, a));
    a.in = n.aa1;
    a.frag = n.fr.f2;
    a.bias = n.fr.b2;
    a.a_out = n.aa2;
    HIP_TRY(launch_conv(32, 2, EPI_HIDDEN, st, a));
    a.in = n.aa2;
    a.frag = n.fr.f3;
    a.bias = n.fr.b3;
    a.a_out = nullptr;
    a.q_old = v0.cur;
    a.q_new = v0.other;
    a.out_scale = out_scale;
    a.rhs = incr ? v0.rhs : nullptr;
    a.scalar_gain = n.input_kind == QSR_SURR_INPUT_STEP ? 1 : 0;
    a.cs = v0.cs;
    a.rows = v0.rows;
    HIP_TRY(launch_conv(32, 1, EPI_APPLY, st, a));
    if (ms_out) {
        HIP_TRY(hipEventRecord(n.ev[1], st));
        HIP_TRY(hipEventSynchronize(n.ev[1]));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, n.ev[0], n.ev[1]));
        *ms_out = ms;
    }
    return QSR_OK;
}

int upload_params(Net& n, const float* host) {
    HIP_TRY(hipSetDevice(n.device));
    HIP_TRY(hipMemcpyAsync(n.P, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, n.st));
    HIP_TRY(hipMemsetAsync(n.M, 0, sizeof(float) * NPARAM, n.st));
    HIP_TRY(hipMemsetAsync(n.Vm, 0, sizeof(float) * NPARAM, n.st));
    HIP_TRY(launch_pack_frags(n.st, n.P, n.fr));
    HIP_TRY(hipStreamSynchronize(n.st));
    n.t = 0;
    n.has_params = true;
    return QSR_OK;
}

}  // namespace

extern "C" {

int qsr_surr_num_params(size_t* n) {
    if (!n) return fail(QSR_E_ARG, "null argument");
    *n = size_t(NPARAM);
    return QSR_OK;
}

int qsr_surr_create(int device, void** out) {
    if (!out) return fail(QSR_E_ARG, "null argument");
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(QSR_E_ARG, "bad device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    auto n = std::make_unique<Net>();
    n->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking));
    for (hipEvent_t& e : n->ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipEventCreateWithFlags(&n->ev_stage, hipEventDisableTiming));
    for (float** p : {&n->P, &n->G, &n->M, &n->Vm}) HIP_TRY(hipMalloc(p, sizeof(float) * NPARAM));
    HIP_TRY(hipMemsetAsync(n->G, 0, sizeof(float) * NPARAM, n->st));
    const long long fe[5] = {frag_elems(8, 2), frag_elems(32, 2), frag_elems(32, 1), frag_elems(8, 2),
                             frag_elems(32, 2)};
    long long bytes = 0;
    for (long long e : fe) bytes += e * 2;
    bytes += (32 + 32 + 16) * 4;
    HIP_TRY(hipMalloc(&n->frag_mem, bytes));
    char* p = static_cast<char*>(n->frag_mem);
    uint16_t** fp[5] = {&n->fr.f1, &n->fr.f2, &n->fr.f3, &n->fr.g3, &n->fr.g2};
    for (int i = 0; i < 5; ++i) {
        *fp[i] = reinterpret_cast<uint16_t*>(p);
        p += fe[i] * 2;
    }
    n->fr.b1 = reinterpret_cast<float*>(p);
    n->fr.b2 = n->fr.b1 + 32;
    n->fr.b3 = n->fr.b2 + 32;
    HIP_TRY(hipStreamSynchronize(n->st));
    *out = n.release();
    return QSR_OK;
}

int qsr_surr_destroy(void* net) {
    Net* n = as_net(net);
    if (!n) return QSR_OK;
    discard(hipSetDevice(n->device));
    discard(hipStreamSynchronize(n->st));
    delete n;
    return QSR_OK;
}

int qsr_surr_init_params(void* net, uint64_t seed) {
    Net* n = as_net(net);
    if (!n) return fail(QSR_E_ARG, "null network");
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases; the last
    // layer scaled by 0.1 so the untrained step starts close to "no change"
    std::vector<float> P(NPARAM);
    SplitMix r{seed};
    auto fill = [&](long long off, long long cnt, int fan_in, double scale) {
        const double a = scale / std::sqrt(double(fan_in));
        for (long long i = 0; i < cnt; ++i) P[off + i] = float((2.0 * r.uniform() - 1.0) * a);
    };
    fill(OFF_W1, OFF_B1 - OFF_W1, TAPS * CH_Q, 1.0);
    fill(OFF_B1, CH_H, TAPS * CH_Q, 1.0);
    fill(OFF_W2, OFF_B2 - OFF_W2, TAPS * CH_H, 1.0);
    fill(OFF_B2, CH_H, TAPS * CH_H, 1.0);
    fill(OFF_W3, OFF_B3 - OFF_W3, TAPS * CH_H, 0.1);
    fill(OFF_B3, CH_Q, TAPS * CH_H, 0.0);
    return upload_params(*n, P.data());
}

int qsr_surr_set_params(void* net, const float* params, size_t count) {
    Net* n = as_net(net);
    if (!n || !params) return fail(QSR_E_ARG, "null argument");
    if ((long long)count != NPARAM) return fail(QSR_E_ARG, "expected " + std::to_string(NPARAM) + " parameters");
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(params[i])) return fail(QSR_E_ARG, "non-finite parameter " + std::to_string(i));
    return upload_params(*n, params);
}

int qsr_surr_get_params(void* net, float* params, size_t count) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!params || (long long)count != NPARAM) return fail(QSR_E_ARG, "bad buffer");
    HIP_TRY(hipSetDevice(n->device));
    HIP_TRY(hipMemcpyAsync(params, n->P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, n->st));
    HIP_TRY(hipStreamSynchronize(n->st));
    return QSR_OK;
}

int qsr_surr_save(void* net, const char* path) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!path) return fail(QSR_E_ARG, "null path");
    std::vector<float> P(NPARAM);
    if (int rc = qsr_surr_get_params(net, P.data(), P.size())) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(QSR_E_ARG, std::string("cannot write ") + path);
    const char magic[8] = {'Q', 'S', 'R', 'S', 'U', 'R', 'R', '1'};
    const int64_t cnt = NPARAM;
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&cnt, sizeof cnt, 1, f) == 1 &&
              std::fwrite(P.data(), sizeof(float), P.size(), f) == P.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(QSR_E_ARG, std::string("short write to ") + path);
    return QSR_OK;
}

int qsr_surr_load(void* net, const char* path) {
    Net* n = as_net(net);
    if (!n || !path) return fail(QSR_E_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(QSR_E_ARG, std::string("cannot read ") + path);
    char magic[8] = {};
    int64_t cnt = 0;
    std::vector<float> P(NPARAM);
    const bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "QSRSURR1", 8) == 0 &&
                    std::fread(&cnt, sizeof cnt, 1, f) == 1 && cnt == NPARAM &&
                    std::fread(P.data(), sizeof(float), P.size(), f) == P.size();
    std::fclose(f);
    if (!ok) return fail(QSR_E_ARG, std::string("not a network file of this layout: ") + path);
    return qsr_surr_set_params(net, P.data(), P.size());
}

}  // extern "C"

namespace {

// One pass over a host batch; with m5 the prediction is m5 (1 + y), else y.
int eval_batch(Net* n, int nbatch, int X, int Y, int Z, const float* x5, const float* m5, const float* target5,
               const float* mask, double* loss_out, float* pred5, float* grad) {
    if (int rc = check_net(n)) return rc;
    if (nbatch < 1 || X < 1 || Y < 1 || Z < 1 || !x5) return fail(QSR_E_ARG, "bad batch");
    if ((loss_out || grad) && Z % 8 != 0) return fail(QSR_E_ARG, "Z must be a multiple of 8 for the loss");
    HIP_TRY(hipSetDevice(n->device));
    const long long V = (long long)nbatch * X * Y * Z;
    if (int rc = ensure_work(*n, V)) return rc;
    if (int rc = ensure_out(*n, 2)) return rc;
    const long long need = V * 16;
    if (need > n->stage_cap) {
        discard(hipFree(n->stage));
        n->stage = nullptr;
        n->stage_cap = 0;
        HIP_TRY(hipMalloc(&n->stage, sizeof(float) * need));
        n->stage_cap = need;
    }
    float* dx = n->stage;
    float* dt = n->stage + V * 5;
    float* dm = n->stage + V * 10;
    float* dg = n->stage + V * 11;
    HIP_TRY(hipMemcpyAsync(dx, x5, sizeof(float) * V * 5, hipMemcpyHostToDevice, n->st));
    if (target5) HIP_TRY(hipMemcpyAsync(dt, target5, sizeof(float) * V * 5, hipMemcpyHostToDevice, n->st));
    if (mask) HIP_TRY(hipMemcpyAsync(dm, mask, sizeof(float) * V, hipMemcpyHostToDevice, n->st));
    if (m5) HIP_TRY(hipMemcpyAsync(dg, m5, sizeof(float) * V * 5, hipMemcpyHostToDevice, n->st));
    n->w.gain = m5 ? (n->input_kind == QSR_SURR_INPUT_STEP ? 2 : 1) : 0;
    HIP_TRY(launch_pack_batch(n->st, dx, target5 ? dt : nullptr, mask ? dm : nullptr, m5 ? dg : nullptr, V, n->w));
    if (int rc = forward(*n, nbatch, X, Y, Z, grad != nullptr)) return rc;
    if (loss_out || grad) {
        if (int rc = loss(*n, n->d_out)) return rc;
        if (grad) {
            if (int rc = backward(*n, nbatch, X, Y, Z)) return rc;
            HIP_TRY(hipMemcpyAsync(grad, n->G, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, n->st));
        }
        double s[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(s, n->d_out, sizeof s, hipMemcpyDeviceToHost, n->st));
        HIP_TRY(hipStreamSynchronize(n->st));
        if (loss_out) *loss_out = s[0] / (5.0 * double(V));
    }
    if (pred5) {
        HIP_TRY(launch_unpack_pred(n->st, n->w, dx));
        HIP_TRY(hipMemcpyAsync(pred5, dx, sizeof(float) * V * 5, hipMemcpyDeviceToHost, n->st));
    }
    HIP_TRY(hipStreamSynchronize(n->st));
    return QSR_OK;
}

}  // namespace

extern "C" {

int qsr_surr_eval(void* net, int nbatch, int X, int Y, int Z, const float* x5, const float* target5,
                  const float* mask, double* loss_out, float* pred5, float* grad) {
    return eval_batch(as_net(net), nbatch, X, Y, Z, x5, nullptr, target5, mask, loss_out, pred5, grad);
}

int qsr_surr_eval_gain(void* net, int nbatch, int X, int Y, int Z, const float* x5, const float* m5,
                       const float* target5, const float* mask, double* loss_out, float* pred5, float* grad) {
    if (!m5) return fail(QSR_E_ARG, "null modulator");
    return eval_batch(as_net(net), nbatch, X, Y, Z, x5, m5, target5, mask, loss_out, pred5, grad);
}

int qsr_surr_adam_step(void* net, const float* grad, double lr) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!grad || !(lr > 0.0) || !std::isfinite(lr)) return fail(QSR_E_ARG, "bad argument");
    HIP_TRY(hipSetDevice(n->device));
    HIP_TRY(hipMemcpyAsync(n->G, grad, sizeof(float) * NPARAM, hipMemcpyHostToDevice, n->st));
    if (int rc = adam(*n, lr)) return rc;
    HIP_TRY(hipStreamSynchronize(n->st));
    return QSR_OK;
}

}  // extern "C"

namespace {

int dataset_add(Net* n, void* ctx, int box, bool with_increment, int* snapshot) {
    if (!n) return fail(QSR_E_ARG, "null network");
    qsrc::FieldView v;
    if (int rc = qsrc::field_view(ctx, box, &v)) return rc;
    if (!v.uploaded) return fail(QSR_E_STATE, "box has no field");
    if (v.device != n->device) return fail(QSR_E_ARG, "network and context are on different devices");
    HIP_TRY(hipSetDevice(n->device));
    double res = 0.0;
    long long cnt = 0;
    if (with_increment) {
        // one explicit step of the whole context into its step buffer; the field is unchanged
        if (int rc = qsrc::euler_residual(ctx, box, &res, &cnt)) return rc;
        if (int rc = qsrc::field_view(ctx, box, &v)) return rc;
    }
    SnapH s;
    s.X = v.box.Nx;
    s.Y = v.box.Ny;
    s.Z = v.box.Nz;
    s.nvox = v.nvox;
    const bool ok = hipMalloc(&s.q, sizeof(float) * 5 * s.nvox) == hipSuccess &&
                    hipMalloc(&s.mask, s.nvox) == hipSuccess &&
                    (!with_increment || hipMalloc(&s.inc, sizeof(float) * 5 * s.nvox) == hipSuccess);
    if (!ok) {
        free_snap(s);
        return fail(QSR_E_HIP, "out of device memory for the snapshot");
    }
    s.inc_rms = cnt > 0 ? res / std::sqrt(5.0 * double(cnt)) : 0.0;
    n->snaps.push_back(s);
    n->d_snaps_stale = true;
    // on the context's stream, ordered after its steps
    HIP_TRY(launch_copy_box(v.st, v.cur, v.cs, v.box.off, s.nvox, s.q));
    HIP_TRY(launch_mask_bytes(v.st, s.nvox, v.box, v.rows, s.mask));
    if (with_increment) HIP_TRY(launch_increment_box(v.st, v.cur, v.rhs, v.cs, v.box, v.rows, s.inc));
    HIP_TRY(hipStreamSynchronize(v.st));
    if (snapshot) *snapshot = int(n->snaps.size()) - 1;
    return QSR_OK;
}

}  // namespace

extern "C" {

int qsr_surr_dataset_add(void* net, void* ctx, int box, int* snapshot) {
    return dataset_add(as_net(net), ctx, box, false, snapshot);
}

int qsr_surr_dataset_add_increment(void* net, void* ctx, int box, int* snapshot) {
    return dataset_add(as_net(net), ctx, box, true, snapshot);
}

int qsr_surr_set_input(void* net, int kind, double horizon) {
    Net* n = as_net(net);
    if (!n) return fail(QSR_E_ARG, "null network");
    if (kind != QSR_SURR_INPUT_FIELD && kind != QSR_SURR_INPUT_INCREMENT && kind != QSR_SURR_INPUT_GAIN &&
        kind != QSR_SURR_INPUT_STEP)
        return fail(QSR_E_ARG, "unknown input kind");
    if (!(horizon > 0.0) || !std::isfinite(horizon)) return fail(QSR_E_ARG, "horizon must be positive");
    n->input_kind = kind;
    n->horizon = horizon;
    return QSR_OK;
}

int qsr_surr_set_tiles(void* net, int mode) {
    Net* n = as_net(net);
    if (!n) return fail(QSR_E_ARG, "null network");
    if (mode != QSR_SURR_TILES_DROPLET && mode != QSR_SURR_TILES_ALL) return fail(QSR_E_ARG, "bad tile mode");
    n->tiles_all = mode == QSR_SURR_TILES_ALL;
    return QSR_OK;
}

int qsr_surr_dataset_size(void* net, int* count) {
    Net* n = as_net(net);
    if (!n || !count) return fail(QSR_E_ARG, "null argument");
    *count = int(n->snaps.size());
    return QSR_OK;
}

int qsr_surr_dataset_clear(void* net) {
    Net* n = as_net(net);
    if (!n) return fail(QSR_E_ARG, "null network");
    HIP_TRY(hipSetDevice(n->device));
    HIP_TRY(hipStreamSynchronize(n->st));
    for (SnapH& s : n->snaps) free_snap(s);
    n->snaps.clear();
    // the device table points at the buffers just freed: drop it, so training
    // rebuilds it even when the next dataset has the same number of snapshots
    discard(hipFree(n->d_snaps));
    n->d_snaps = nullptr;
    n->d_snaps_stale = true;
    return QSR_OK;
}

int qsr_surr_train(void* net, const int* pairs, int npairs, int iters, int batch, int crop, double lr, uint64_t seed,
                   double weight_floor, double* loss_per_iter) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!pairs || npairs < 1 || iters < 1 || batch < 1 || !(lr > 0.0) || !std::isfinite(lr) ||
        !(weight_floor > 0.0) || !std::isfinite(weight_floor))
        return fail(QSR_E_ARG, "bad argument");
    if (crop < 8 || crop % 8 != 0) return fail(QSR_E_ARG, "crop must be a positive multiple of 8");
    const int ns = int(n->snaps.size());
    for (int p = 0; p < npairs; ++p) {
        const int a = pairs[2 * p], b = pairs[2 * p + 1];
        if (a < 0 || a >= ns || b < 0 || b >= ns) return fail(QSR_E_ARG, "pair " + std::to_string(p) + ": bad snapshot");
        const SnapH &sa = n->snaps[a], &sb = n->snaps[b];
        if (sa.X != sb.X || sa.Y != sb.Y || sa.Z != sb.Z)
            return fail(QSR_E_ARG, "pair " + std::to_string(p) + ": snapshots differ in size");
        if (sa.X < crop || sa.Y < crop || sa.Z < crop)
            return fail(QSR_E_ARG, "pair " + std::to_string(p) + ": box smaller than the crop");
        if (n->input_kind != QSR_SURR_INPUT_FIELD && !(sa.inc && sa.inc_rms > 0.0 && std::isfinite(sa.inc_rms)))
            return fail(QSR_E_ARG, "pair " + std::to_string(p) + ": input snapshot has no (nonzero) increment");
    }
    HIP_TRY(hipSetDevice(n->device));
    const long long V = (long long)batch * crop * crop * crop;
    if (int rc = ensure_work(*n, V)) return rc;
    if (int rc = ensure_out(*n, 2LL * std::max(iters, npairs))) return rc;
    if (n->d_snaps_stale) {
        discard(hipFree(n->d_snaps));
        n->d_snaps = nullptr;
        std::vector<Snap> hs(n->snaps.size());
        for (size_t i = 0; i < hs.size(); ++i) {
            const SnapH& s = n->snaps[i];
            hs[i] = Snap{s.q, s.mask, s.X, s.Y, s.Z, 0, s.nvox, s.inc};
        }
        HIP_TRY(hipMalloc(&n->d_snaps, sizeof(Snap) * hs.size()));
        HIP_TRY(hipMemcpy(n->d_snaps, hs.data(), sizeof(Snap) * hs.size(), hipMemcpyHostToDevice));
        n->d_snaps_stale = false;
    }
    const long long njobs = (long long)iters * batch;
    if (njobs > n->jobs_cap) {
        discard(hipFree(n->d_jobs));
        n->d_jobs = nullptr;
        n->jobs_cap = 0;
        HIP_TRY(hipMalloc(&n->d_jobs, sizeof(CropJob) * njobs));
        n->jobs_cap = njobs;
    }
    // Each pair's scored voxels are weighted by 1 / max(ms, weight_floor), ms
    // the mean square of its target: pairs with small changes (late in a
    // relaxation) count more, up to the floor.  A floor >= 1 weights all alike.
    for (int p = 0; p < npairs; ++p) {
        const SnapH &sa = n->snaps[pairs[2 * p]], &sb = n->snaps[pairs[2 * p + 1]];
        long long nblk = 0;
        HIP_TRY(launch_pair_power(n->st, sa.q, sb.q, sa.mask, sa.nvox, n->w.red, &nblk));
        HIP_TRY(launch_sum_slots(n->st, n->w.red, nblk, 2, n->d_out + 2LL * p));
    }
    std::vector<double> pw(2 * size_t(npairs));
    HIP_TRY(hipMemcpyAsync(pw.data(), n->d_out, sizeof(double) * pw.size(), hipMemcpyDeviceToHost, n->st));
    HIP_TRY(hipStreamSynchronize(n->st));
    // increment and gain input: target = (q_out - q_in) / (N s), modulator inc / s, s of the
    // input snapshot; the network reads inc / s (increment) or q_in (gain)
    const bool incr = n->input_kind != QSR_SURR_INPUT_FIELD;
    std::vector<float> weight(npairs, 1.0f), in_scale(npairs, 1.0f), out_scale(npairs, 1.0f);
    for (int p = 0; p < npairs; ++p) {
        if (incr) {
            const double s_in = n->snaps[pairs[2 * p]].inc_rms;
            in_scale[p] = float(1.0 / s_in);
            out_scale[p] = float(1.0 / (n->horizon * s_in));
        }
        const double ms = pw[2 * p + 1] > 0 ? pw[2 * p] / (5.0 * pw[2 * p + 1]) : 0.0;
        weight[p] = float(1.0 / std::max(ms * double(out_scale[p]) * double(out_scale[p]), weight_floor));
    }
    std::vector<CropJob> jobs(njobs);
    SplitMix r{seed};
    for (long long j = 0; j < njobs; ++j) {
        const int p = r.below(npairs);
        const SnapH& s = n->snaps[pairs[2 * p]];
        const int ox = r.below(s.X - crop + 1), oy = r.below(s.Y - crop + 1), oz = r.below(s.Z - crop + 1);
        jobs[j] = CropJob{pairs[2 * p], pairs[2 * p + 1], ox, oy, oz, weight[p], n->input_kind, in_scale[p],
                          out_scale[p], 0};
    }
    HIP_TRY(hipStreamSynchronize(n->st));
    HIP_TRY(hipMemcpy(n->d_jobs, jobs.data(), sizeof(CropJob) * njobs, hipMemcpyHostToDevice));
    n->w.gain = n->input_kind == QSR_SURR_INPUT_STEP ? 2 : (incr ? 1 : 0);
    for (int it = 0; it < iters; ++it) {
        HIP_TRY(launch_crop(n->st, n->d_snaps, n->d_jobs + (long long)it * batch, batch, crop, n->w));
        if (int rc = forward(*n, batch, crop, crop, crop, true)) return rc;
        if (int rc = loss(*n, n->d_out + 2LL * it)) return rc;
        if (int rc = backward(*n, batch, crop, crop, crop)) return rc;
        if (int rc = adam(*n, lr)) return rc;
    }
    std::vector<double> h(2 * size_t(iters));
    HIP_TRY(hipMemcpyAsync(h.data(), n->d_out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, n->st));
    HIP_TRY(hipStreamSynchronize(n->st));
    if (loss_per_iter)
        for (int it = 0; it < iters; ++it)
            loss_per_iter[it] = h[2 * it + 1] > 0 ? h[2 * it] / (5.0 * h[2 * it + 1]) : 0.0;
    return QSR_OK;
}

int qsr_surr_apply_ladder(void* net, void* ctx, int mode, const double* scales, int nscales,
                          qsr_surr_result* out) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!out || mode < QSR_SURR_CHECKED || mode > QSR_SURR_PROPOSE || !scales || nscales < 1)
        return fail(QSR_E_ARG, "bad argument");
    for (int i = 0; i < nscales; ++i)
        if (!(scales[i] > 0.0) || !std::isfinite(scales[i])) return fail(QSR_E_ARG, "scales must be positive");
    if (mode != QSR_SURR_CHECKED) nscales = 1;
    qsrc::FieldView v;
    if (int rc = qsrc::field_view(ctx, 0, &v)) return rc;
    if (v.nboxes != 1) return fail(QSR_E_ARG, "a batched context takes qsr_surr_apply_boxes");
    if (!v.uploaded) return fail(QSR_E_STATE, "box has no field");
    if (!v.active) return fail(QSR_E_STATE, "box is held");
    if (v.device != n->device) return fail(QSR_E_ARG, "network and context are on different devices");
    HIP_TRY(hipSetDevice(n->device));
    HIP_TRY(hipStreamSynchronize(n->st));  // parameters last written on the network's stream
    qsr_surr_result r{};
    const qsr_obs_opts oo = {-1, 0.1, 0.25, -1, -1.0, 0, 0, 1};  // energies only
    qsr_obs o{};
    if (int rc = qsr_observables(ctx, 0, &oo, &o)) return rc;
    r.F_before = o.total;
    long long cnt = 0;
    if (int rc = qsrc::euler_residual(ctx, 0, &r.res_before, &cnt)) return rc;
    if (int rc = qsrc::field_view(ctx, 0, &v)) return rc;  // the step buffer now holds E(Q)

    // increment and gain input modulate u = E(Q) - Q (s its RMS per component): nothing to
    // propose if it is zero
    const bool incr = n->input_kind != QSR_SURR_INPUT_FIELD;
    const double s_inc = cnt > 0 ? r.res_before / std::sqrt(5.0 * double(cnt)) : 0.0;
    if (incr && !(s_inc > 0.0 && std::isfinite(s_inc))) {
        r.F_after = r.F_before;
        r.res_after = r.res_before;
        *out = r;
        return QSR_OK;
    }

    long long nv = 0;
    if (int rc = propose_list(*n, {Prop{v, s_inc}}, float(scales[0] * (incr ? n->horizon : 1.0)), &r.ms_network,
                              &nv))
        return rc;
    r.net_voxels = nv;
    const hipStream_t st = v.st;

    // proposal i lies in the other buffer; the previous field stays in the current one
    bool keep = false;
    for (int i = 0; i < nscales && !keep; ++i) {
        if (i > 0) {
            if (int rc = qsrc::field_view(ctx, 0, &v)) return rc;
            HIP_TRY(launch_rescale_step(st, v.cur, v.other, v.cs, v.box, v.rows, float(scales[i] / scales[i - 1])));
        }
        r.scale = scales[i];
        r.tried = i + 1;
        if (int rc = qsrc::commit_other(ctx)) return rc;
        if (int rc = qsr_observables(ctx, 0, &oo, &o)) return rc;
        r.F_after = o.total;
        if (int rc = qsrc::euler_residual(ctx, 0, &r.res_after)) return rc;
        if (mode == QSR_SURR_FORCE) keep = true;
        else if (mode == QSR_SURR_CHECKED)
            keep = std::isfinite(r.F_after) && std::isfinite(r.res_after) && r.F_after <= r.F_before &&
                   r.res_after <= r.res_before;
        if (!keep) {
            // back to the previous field; the proposal stays in the other buffer
            if (int rc = qsrc::commit_other(ctx)) return rc;
        }
    }
    r.accepted = keep ? 1 : 0;
    *out = r;
    return QSR_OK;
}

int qsr_surr_apply(void* net, void* ctx, int mode, qsr_surr_result* out) {
    const double one = 1.0;
    return qsr_surr_apply_ladder(net, ctx, mode, &one, 1, out);
}

int qsr_surr_apply_boxes(void* net, void* ctx, int mode, const double* scales, int nscales, const int* boxes,
                         int nsel, qsr_surr_result* out) {
    Net* n = as_net(net);
    if (int rc = check_net(n)) return rc;
    if (!out || !boxes || nsel < 1 || mode < QSR_SURR_CHECKED || mode > QSR_SURR_PROPOSE || !scales || nscales < 1)
        return fail(QSR_E_ARG, "bad argument");
    for (int i = 0; i < nscales; ++i)
        if (!(scales[i] > 0.0) || !std::isfinite(scales[i])) return fail(QSR_E_ARG, "scales must be positive");
    if (mode != QSR_SURR_CHECKED) nscales = 1;
    qsrc::FieldView v;
    if (int rc = qsrc::field_view(ctx, 0, &v)) return rc;
    if (v.device != n->device) return fail(QSR_E_ARG, "network and context are on different devices");
    std::vector<char> seen(static_cast<size_t>(v.nboxes), 0);
    for (int i = 0; i < nsel; ++i) {
        const int b = boxes[i];
        if (b < 0 || b >= v.nboxes) return fail(QSR_E_ARG, "box " + std::to_string(b) + " out of range");
        if (seen[b]) return fail(QSR_E_ARG, "box " + std::to_string(b) + " listed twice");
        seen[b] = 1;
        if (int rc = qsrc::field_view(ctx, b, &v)) return rc;
        if (!v.uploaded) return fail(QSR_E_STATE, "box " + std::to_string(b) + " has no field");
        if (!v.active) return fail(QSR_E_STATE, "box " + std::to_string(b) + " is held");
    }
    HIP_TRY(hipSetDevice(n->device));
    HIP_TRY(hipStreamSynchronize(n->st));  // parameters last written on the network's stream
    const bool incr = n->input_kind != QSR_SURR_INPUT_FIELD;
    std::vector<qsr_surr_result> R(static_cast<size_t>(nsel));
    std::vector<double> s_inc(static_cast<size_t>(nsel), 0.0), F(static_cast<size_t>(nsel)),
        res(static_cast<size_t>(nsel));
    std::vector<long long> cnt(static_cast<size_t>(nsel), 0);
    // energies and residuals before, from one explicit step of every active box
    if (int rc = qsrc::euler_rhs(ctx)) return rc;
    if (int rc = qsr_energies(ctx, boxes, nsel, F.data())) return rc;
    if (int rc = qsrc::box_residuals(ctx, boxes, nsel, res.data(), cnt.data())) return rc;
    for (int i = 0; i < nsel; ++i) {
        R[i].F_before = R[i].F_after = F[i];
        R[i].res_before = R[i].res_after = res[i];
        s_inc[i] = cnt[i] > 0 ? res[i] / std::sqrt(5.0 * double(cnt[i])) : 0.0;
    }
    // proposals into the other buffer, which holds a copy of the current field elsewhere
    if (int rc = qsrc::copy_current_to_other(ctx)) return rc;
    std::vector<int> pending;
    std::vector<Prop> props;
    for (int i = 0; i < nsel; ++i) {
        if (incr && !(s_inc[i] > 0.0 && std::isfinite(s_inc[i]))) continue;  // nothing to propose
        if (int rc = qsrc::field_view(ctx, boxes[i], &v)) return rc;
        props.push_back(Prop{v, s_inc[i]});
        pending.push_back(i);
    }
    double phase_ms = 0.0;  // the device time of all listed boxes' passes
    long long nv = 0;
    if (int rc = propose_list(*n, props, float(scales[0] * (incr ? n->horizon : 1.0)), &phase_ms, &nv)) return rc;
    for (int i : pending) {
        R[i].ms_network = phase_ms;
        R[i].net_voxels = nv;
    }
    if (!pending.empty()) {
        // current: the proposals (every other box unchanged); other: the previous field
        if (int rc = qsrc::swap_buffers(ctx)) return rc;
        for (int r = 0; r < nscales && !pending.empty(); ++r) {
            if (r > 0)
                for (int i : pending) {
                    if (int rc = qsrc::field_view(ctx, boxes[i], &v)) return rc;
                    // the context owns both buffers: the current one is rescaled in place
                    HIP_TRY(launch_rescale_step(v.st, v.other, const_cast<float*>(v.cur), v.cs, v.box, v.rows,
                                                float(scales[r] / scales[r - 1])));
                }
            if (int rc = qsrc::euler_rhs(ctx)) return rc;
            std::vector<int> pb;
            for (int i : pending) pb.push_back(boxes[i]);
            if (int rc = qsr_energies(ctx, pb.data(), int(pb.size()), F.data())) return rc;
            if (int rc = qsrc::box_residuals(ctx, pb.data(), int(pb.size()), res.data(), nullptr)) return rc;
            std::vector<int> still;
            for (size_t m = 0; m < pending.size(); ++m) {
                const int i = pending[m];
                R[i].scale = scales[r];
                R[i].tried = r + 1;
                R[i].F_after = F[m];
                R[i].res_after = res[m];
                bool keep = mode == QSR_SURR_FORCE;
                if (mode == QSR_SURR_CHECKED)
                    keep = std::isfinite(R[i].F_after) && std::isfinite(R[i].res_after) &&
                           R[i].F_after <= R[i].F_before && R[i].res_after <= R[i].res_before;
                if (keep) R[i].accepted = 1;
                else still.push_back(i);
            }
            pending.swap(still);
        }
        for (int i : pending)  // refused at every scale: back to the previous field
            if (int rc = qsrc::copy_box_other_to_current(ctx, boxes[i])) return rc;
        if (int rc = qsrc::field_view(ctx, 0, &v)) return rc;
        HIP_TRY(hipStreamSynchronize(v.st));
    }
    for (int i = 0; i < nsel; ++i) out[i] = R[i];
    return QSR_OK;
}

}  // extern "C"
