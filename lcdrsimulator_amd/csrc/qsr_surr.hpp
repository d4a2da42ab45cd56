// qsr_surr.hpp -- device-side layout of the learned warm-start step and the
// host launchers of its kernels (qsr_surr.hip).
//
// Network: three 3x3x3 "same" convolutions (zero padding), 5 -> 32 -> 32 -> 5
// channels, GELU (tanh form) after the first two.  It maps a five-component
// field (the independent components Qxx, Qxy, Qxz, Qyy, Qyz of Q itself, or of
// one explicit step's increment of Q, normalised) to a predicted change dQ.
//
//   z1 = conv(W1, x) + b1     a1 = bf16(gelu(z1))
//   z2 = conv(W2, a1) + b2    a2 = bf16(gelu(z2))
//   y  = conv(W3, a2) + b3    (FP32)
//   conv(W, a)[v, co] = sum_{t, ci} W[t][ci][co] * a[v + off(t), ci]
//   t = (di+1)*9 + (dj+1)*3 + (dk+1), off(t) = (di, dj, dk) along (i, j, k)
//
// Operands are BF16 (x, W, a1, a2, and in the backward pass the output
// gradients), every product is accumulated in FP32 on the matrix cores.
// Master parameters, gradients and the Adam moments are FP32, in one flat
// array: W1[27][5][32] b1[32] W2[27][32][32] b2[32] W3[27][32][5] b3[5].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "qsr_kernels.hpp"

namespace qsrs {

constexpr int TAPS = 27;
constexpr int CH_Q = 5, CH_H = 32;
constexpr long long OFF_W1 = 0;
constexpr long long OFF_B1 = OFF_W1 + TAPS * CH_Q * CH_H;
constexpr long long OFF_W2 = OFF_B1 + CH_H;
constexpr long long OFF_B2 = OFF_W2 + TAPS * CH_H * CH_H;
constexpr long long OFF_W3 = OFF_B2 + CH_H;
constexpr long long OFF_B3 = OFF_W3 + TAPS * CH_H * CH_Q;
constexpr long long NPARAM = OFF_B3 + CH_Q;  // 36,357

// Activation tensors are voxel-major with the channels innermost ("NDHWC"),
// padded to CP = 8 (the Q field and its gradient) or 32 channels, over a batch
// of n boxes of X x Y x Z voxels: element (v, c) at v * CP + c, with
// v = ((b * X + i) * Y + j) * Z + k.  The weight-gradient kernel also reads
// channel-major copies ("cm"): element (v, c) at c * V + v.
//
// Weight fragments: a convolution with CP input channels is a GEMM with
// K = 27 * CP (tap-major, channel fastest) padded to a multiple of 32, cut
// into KSTEPS = K / 32 steps.  For output block nb (16 channels) and step ks,
// lane l of a wave holds the 8 BF16 values W[K = ks*32 + (l/16)*8 + e][co = nb*16 + l%16],
// e = 0..7: the A operand of v_mfma_f32_16x16x32_bf16 with the output channels as M.
constexpr int ksteps(int cp) { return (TAPS * cp + 31) / 32; }
constexpr long long frag_elems(int cp, int nb) { return (long long)nb * ksteps(cp) * 64 * 8; }

struct Frags {
    uint16_t* f1;  // layer 1 forward: CP 8, 2 blocks
    uint16_t* f2;  // layer 2 forward: CP 32, 2 blocks
    uint16_t* f3;  // layer 3 forward: CP 32, 1 block (5 of 16 outputs used)
    uint16_t* g3;  // layer 3 input gradient: CP 8 (dy), 2 blocks (da2)
    uint16_t* g2;  // layer 2 input gradient: CP 32 (dz2), 2 blocks (da1)
    float* b1;     // 32
    float* b2;     // 32
    float* b3;     // 16 (5 used)
};

// Convolution epilogues.
enum : int {
    EPI_HIDDEN = 0,  // z = acc + b -> optional FP32 z, a = bf16(gelu(z)) (NDHWC and optional cm)
    EPI_PRED = 1,    // y = acc + b -> FP32 NDHWC, 8 channels (5 used)
    EPI_APPLY = 2,   // q_new = q_old + out_scale * y (field input) or
                     // q_old + out_scale * u (1 + y), u = rhs - q_old (increment input),
                     // on the voxels a step can change
    EPI_BWD = 3,     // dz = bf16(acc * gelu'(z_in)) (NDHWC and cm)
};

// One box of a tile-listed convolution launch: its dimensions, the index of its
// first voxel in the launch's packed activation tensors, and (apply epilogue)
// its description in the context.
struct ConvBox {
    int X, Y, Z, pad;
    long long act_off;
    qsrk::BoxDev box;
};

// Tiles of a tile-listed launch: TILE_TJ rows (one wave each) x TILE_TK voxels of
// k x the launch's planes per tile.  A box's list holds every tile with a voxel
// within TILE_REACH of one a step can change: the three layers' outputs there
// only read inputs the list computes, so the network's output on the voxels a
// step can change is the same as over the whole box.
constexpr int TILE_TJ = 12, TILE_TK = 32, TILE_REACH = 2;

struct ConvArgs {
    const uint16_t* in;     // NDHWC BF16, CP channels
    const uint16_t* frag;   // weight fragments
    const float* bias;      // forward epilogues
    int n, X, Y, Z;
    long long V, nseg;      // voxels, 16-voxel k segments
    int segs_row;
    float* z_out;           // EPI_HIDDEN (optional)
    uint16_t* a_out;        // EPI_HIDDEN / EPI_BWD, NDHWC 32
    uint16_t* a_cm;         // EPI_HIDDEN / EPI_BWD (optional)
    const float* z_in;      // EPI_BWD
    float* pred;            // EPI_PRED
    const float* q_old;     // EPI_APPLY: SoA planes of the context
    float* q_new;
    float out_scale;
    const float* rhs;       // EPI_APPLY under increment input: one explicit step of q_old (else null)
    int scalar_gain;        // EPI_APPLY with rhs: u (1 + y_0) for every component (step-size input)
    long long cs;
    qsrk::BoxDev box;
    const qsrk::RowMask* rows;
    // tile-listed launch (null for a uniform batch): block b computes tiles[b] = (index
    // into boxes, i0, j0, k0) over tile_ch planes; in and the activation outputs are
    // packed box after box (ConvBox::act_off), and the apply epilogue takes each tile's
    // box from boxes instead of `box`
    const int4* tiles;
    const ConvBox* boxes;
    int ntiles, tile_ch;
};

// One dataset snapshot: five FP32 planes of nvox voxels plus a byte per voxel
// (1 = a plain step can change it), and optionally the increment of one
// explicit step from it in the same layout (null if not recorded).
struct Snap {
    const float* q;
    const uint8_t* mask;
    int X, Y, Z, pad;
    long long nvox;
    const float* inc;
};
// One training crop: pair (input snapshot, target snapshot), its origin, the
// weight of its scored voxels (the inverse of the pair's target mean square)
// and the input kind.  Field input: x = q_in, target = q_out - q_in.
// Increment and step-size input: x = inc_in * in_scale, target = (q_out - q_in) *
// out_scale, modulator inc_in * in_scale.  Gain input: x = q_in, the rest as increment input.
struct CropJob {
    int s_in, s_out, ox, oy, oz;
    float weight;
    int kind;
    float in_scale, out_scale;
    int pad;
};

// Training workspace views (batch of n crops of edge C, V = n*C^3 voxels).
struct Work {
    long long V;
    uint16_t *x, *x_cm;          // 8 ch
    float *z1, *z2;              // 32 ch
    uint16_t *a1, *a1_cm, *a2, *a2_cm;
    float *pred, *tgt, *mask;    // 8 ch, 8 ch, 1
    float* mod;                  // 8 ch FP32: the modulator m of a modulated prediction
    int gain;                    // 1: the prediction is m (1 + y); 2: m (1 + y_0) (one gain per voxel); 0: y
    uint16_t *dy, *dy_cm;        // 8 ch
    uint16_t *dz2, *dz2_cm, *dz1, *dz1_cm;
    float* part;                 // weight-gradient partials
    double* red;                 // loss partials
};

constexpr int WG_BLOCKS = 256;  // weight-gradient kernel grid (partials per layer)
constexpr long long part_floats() { return (long long)WG_BLOCKS * 880 * 32; }

hipError_t launch_conv(int cp, int nb, int epi, hipStream_t st, const ConvArgs& a);
hipError_t launch_wgrad(int cp, int nb, hipStream_t st, const uint16_t* act_cm, const uint16_t* dz_cm, int X, int Y,
                        int Z, long long V, float* part);
// Sums the partials of one layer into the flat gradient (weights then bias).
hipError_t launch_wgrad_reduce(int cp, int nb, int cin_real, int cout_real, long long w_off, long long b_off,
                               hipStream_t st, const float* part, float* grad);
hipError_t launch_pack_frags(hipStream_t st, const float* params, Frags f);
hipError_t launch_adam(hipStream_t st, float* params, const float* grad, float* m, float* v, long long n, float lr_t,
                       float b1, float b2, float eps, float inv_sqrt_bc2);
hipError_t launch_crop(hipStream_t st, const Snap* snaps, const CropJob* jobs, int nbatch, int C, const Work& w);
// Host batch (staged on the device): x5, t5 (may be null), mask (may be null), and the
// modulator m5 of a modulated prediction (may be null).
hipError_t launch_pack_batch(hipStream_t st, const float* x5, const float* t5, const float* m, const float* m5,
                             long long V, const Work& w);
hipError_t launch_pack_field(hipStream_t st, const float* q, long long cs, long long off, long long V, uint16_t* x);
// Output gradient of the masked mean-square loss: dy = 2 m (y - t) / (5 V) in BF16;
// partials of sum m (y - t)^2 and sum m per block; returns the block count.
hipError_t launch_loss(hipStream_t st, const Work& w, float scale, long long* nblk);
hipError_t launch_sum_slots(hipStream_t st, const double* partials, long long nblk, int nslot, double* out);
// Partials (per block: sum over the voxels a step can change of |b - a|^2, their
// count) of two snapshots of one box; returns the block count.
hipError_t launch_pair_power(hipStream_t st, const float* qa, const float* qb, const uint8_t* mask, long long nvox,
                             double* partials, long long* nblk);
hipError_t launch_mask_bytes(hipStream_t st, long long nvox, const qsrk::BoxDev& box, const qsrk::RowMask* rows,
                             uint8_t* out);
// Partials (per block: sum of |rhs - q|^2, voxel count) over the voxels a step can
// change, box faces excluded.
hipError_t launch_residual(hipStream_t st, const float* q, const float* rhs, long long cs, const qsrk::BoxDev& box,
                           const qsrk::RowMask* rows, double* partials, long long* nblk);
// Blocks (= partial records) of launch_residual's grid for a box of nvox voxels.
long long residual_blocks(long long nvox);
// The residual partials and their sums (out + 2 m: sum |rhs - q|^2, count) of n boxes in two
// launches: box m is boxes[list[m]] and owns partial records [prefix[m], prefix[m+1]), prefix[m+1] -
// prefix[m] = residual_blocks(its voxels), total_blocks = prefix[n]; each box's values equal
// launch_residual + launch_sum_slots on it alone, bit for bit.
hipError_t launch_residual_multi(hipStream_t st, const float* q, const float* rhs, long long cs,
                                 const qsrk::BoxDev* boxes, const int* list, const long long* prefix, int n,
                                 long long total_blocks, const qsrk::RowMask* rows, double* partials, double* out);
hipError_t launch_copy_box(hipStream_t st, const float* q, long long cs, long long off, long long nvox, float* dst);
// One explicit step's increment rhs - q of a box on the voxels the residual sums
// over (zero elsewhere), into five planes of the box's voxel count.
hipError_t launch_increment_box(hipStream_t st, const float* q, const float* rhs, long long cs,
                                const qsrk::BoxDev& box, const qsrk::RowMask* rows, float* dst);
// Network input (NDHWC, 8 channels, BF16) from the increment scaled by inv_s.
hipError_t launch_pack_increment(hipStream_t st, const float* q, const float* rhs, long long cs,
                                 const qsrk::BoxDev& box, const qsrk::RowMask* rows, float inv_s, uint16_t* x);
// other = cur + ratio (other - cur) on the voxels a step can change.
hipError_t launch_rescale_step(hipStream_t st, const float* cur, float* other, long long cs, const qsrk::BoxDev& box,
                               const qsrk::RowMask* rows, float ratio);
// (nbatch, X, Y, Z, 5) output: y, or m (1 + y) (w.gain 1), or m (1 + y_0) (w.gain 2).
hipError_t launch_unpack_pred(hipStream_t st, const Work& w, float* out5);

}  // namespace qsrs
