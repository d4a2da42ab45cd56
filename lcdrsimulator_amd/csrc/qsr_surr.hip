// qsr_surr.hip -- kernels of the learned warm-start step: BF16 implicit-GEMM
// 3x3x3 convolutions on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation), their input and weight gradients, the masked loss, Adam
// and the packing of parameters into MFMA operand fragments.  The network and
// the layouts are described in qsr_surr.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "qsr_surr.hpp"

namespace {

using qsrk::BoxDev;
using qsrk::RowMask;
using namespace qsrs;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr float kGeluC = 0.7978845608028654f;  // sqrt(2/pi)
constexpr float kGeluA = 0.044715f;

// GELU, tanh form: gelu(z) = z (1 + tanh u) / 2 with u = c (z + a z^3).  Since
// (1 + tanh u) / 2 = 1 / (1 + e^(-2u)), it is evaluated as z / (1 + 2^(-2u log2 e))
// with the hardware exp2 and reciprocal (e^(-2u) -> inf gives 0, -> 0 gives z).
constexpr float kGeluK = -2.0f * kGeluC * 1.4426950408889634f;  // -2 c log2(e)
__device__ __forceinline__ float gelu_sig(float z, float z2) {
    const float e = __builtin_amdgcn_exp2f(kGeluK * (z * __builtin_fmaf(kGeluA, z2, 1.0f)));
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu_f(float z) { return z * gelu_sig(z, z * z); }
// d/dz of gelu_f: s + 2 z s (1 - s) u'(z), s = 1 / (1 + e^(-2u)), u' = c (1 + 3 a z^2)
__device__ __forceinline__ float gelu_d(float z) {
    const float z2 = z * z;
    const float sg = gelu_sig(z, z2);
    return __builtin_fmaf(2.0f * kGeluC * z * sg * (1.0f - sg), __builtin_fmaf(3.0f * kGeluA, z2, 1.0f), sg);
}

// Round to nearest even (the hardware conversion), as raw bits.
__device__ __forceinline__ uint16_t bf16_bits(float x) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
}

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
}

// Workgroup barrier ordering LDS only: a full __syncthreads() also waits for
// every outstanding global access, including the output stores in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ bool changeable(const BoxDev& B, const RowMask* rows, int i, int j, int k) {
    if (B.flags & qsrk::F_PERIODIC) return true;
    const RowMask m = rows[B.rows_off + (long long)i * B.Ny + j];
    const bool evolve = k >= m.v[0] && k <= m.v[1];
    const bool shell = (k >= m.v[2] && k <= m.v[3]) && !(k >= m.v[4] && k <= m.v[5]);
    return evolve && !(shell && (B.flags & (qsrk::F_SKIPSHELL | qsrk::F_BC)));
}

// The voxels the explicit-Euler residual sums over: changeable, box faces excluded.
__device__ __forceinline__ bool stepped_interior(const BoxDev& B, const RowMask* rows, int i, int j, int k) {
    const bool face = i == 0 || j == 0 || k == 0 || i == B.Nx - 1 || j == B.Ny - 1 || k == B.Nz - 1;
    return !face && changeable(B, rows, i, j, k);
}

// ---------------------------------------------------------------------------
// Convolution (forward, and the input gradient as a forward convolution with
// the taps flipped and the channel roles swapped) as an implicit GEMM on the
// matrix cores:  C[co][voxel] += Wfrag[co][K] * act[K][voxel], K = (tap, ci)
// in steps of 32.
//
// A block owns TJ rows (j, one wave each) x 16*S voxels (k) and streams along i over `ch`
// planes of one batch element.  Its LDS holds the layer's weight fragments
// and a ring of three activation planes (i-1, i, i+1) with a one-voxel halo in
// j and k (zero outside the box), so all 27 taps of a plane are LDS reads; the
// global loads of plane i+2 are issued before the matrix work of plane i and
// land in the ring after it.  Wave w computes row j0+w, S segments of 16
// voxels, all 16*NB output channels.  LDS plane layout: [row][channel group of
// 8][k] in 16-byte units, so the 16 lanes of a group read consecutive units.
// ---------------------------------------------------------------------------
template <int CP, int S, int TJ_>
struct ConvTile {
    static constexpr int G = CP / 8;          // 16-byte channel groups per voxel
    static constexpr int TK = 16 * S, TJ = TJ_, NT = 64 * TJ_;
    static constexpr int RW = TK + 2, RR = TJ + 2;
    static constexpr int SLOT = RR * G * RW;  // 16-byte units per plane
    static constexpr int UPT = (SLOT + NT - 1) / NT;
};

// LDS of a convolution launch in 16-byte units: weight fragments, the three plane slots, and
// for the apply epilogue a per-wave 5 x 16S float transpose area.
template <int CP, int NB, int EPI, int S, int TJ>
constexpr int conv_lds_units() {
    return NB * ksteps(CP) * 64 + 3 * ConvTile<CP, S, TJ>::SLOT + (EPI == EPI_APPLY ? TJ * CH_Q * 16 * S / 4 : 0);
}

template <int CP, int NB, int EPI, int S, int TJ>
__global__ __launch_bounds__(64 * TJ) void conv_kernel(ConvArgs a, int ch, int ntj, int ntk, int nti) {
    using T = ConvTile<CP, S, TJ>;
    constexpr int KS = ksteps(CP);
    constexpr int G = T::G, RW = T::RW, SLOT = T::SLOT, UPT = T::UPT, NT = T::NT;
    extern __shared__ u32x4 smem[];
    u32x4* wl = smem;                    // NB * KS * 64 weight fragments
    u32x4* ring = smem + NB * KS * 64;   // 3 plane slots
    float* ytr = reinterpret_cast<float*>(ring + 3 * SLOT);  // apply epilogue: 5 x 16S floats per wave
    const u32x4* wf = reinterpret_cast<const u32x4*>(a.frag);
    for (int f = threadIdx.x; f < NB * KS * 64; f += NT) wl[f] = wf[f];

    // the block's tile: an entry of the tile list, or decoded from the block index over the
    // uniform batch; bbase is the first voxel of its box (batch element) in the activations
    int X = a.X, Y = a.Y, Z = a.Z, i0, j0, k0;
    long long bbase;
    BoxDev B = a.box;
    if (a.tiles) {
        const int4 t4 = a.tiles[blockIdx.x];
        const ConvBox* cb = a.boxes + t4.x;
        X = cb->X;
        Y = cb->Y;
        Z = cb->Z;
        bbase = cb->act_off;
        if constexpr (EPI == EPI_APPLY) B = cb->box;
        i0 = t4.y;
        j0 = t4.z;
        k0 = t4.w;
    } else {
        int bid = blockIdx.x;
        const int kt = bid % ntk;
        bid /= ntk;
        const int jt = bid % ntj;
        bid /= ntj;
        const int it = bid % nti;
        const int b = bid / nti;
        k0 = kt * T::TK;
        j0 = jt * T::TJ;
        i0 = it * ch;
        bbase = (long long)b * X * Y * Z;
    }
    const int i1 = min(i0 + ch, X);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, col = lane & 15;
    const int j = j0 + w;
    const long long YZ = (long long)Y * Z;

    // staging: unit u = (r * RW + kk) * G + gg, consecutive threads read consecutive 16 B
    auto load_plane = [&](int ip, u32x4 (&st)[UPT]) {
#pragma unroll
        for (int s = 0; s < UPT; ++s) {
            const int u = threadIdx.x + s * NT;
            st[s] = u32x4{0u, 0u, 0u, 0u};
            if (u < SLOT) {
                const int gg = u % G, kk = (u / G) % RW, r = u / (G * RW);
                const int jj = j0 - 1 + r, kq = k0 - 1 + kk;
                if (unsigned(ip) < unsigned(X) && unsigned(jj) < unsigned(Y) && unsigned(kq) < unsigned(Z))
                    st[s] = *reinterpret_cast<const u32x4*>(
                        a.in + (bbase + (long long)ip * YZ + (long long)jj * Z + kq) * CP + gg * 8);
            }
        }
    };
    auto store_plane = [&](int slot, const u32x4 (&st)[UPT]) {
#pragma unroll
        for (int s = 0; s < UPT; ++s) {
            const int u = threadIdx.x + s * NT;
            if (u < SLOT) {
                const int gg = u % G, kk = (u / G) % RW, r = u / (G * RW);
                ring[slot * SLOT + (r * G + gg) * RW + kk] = st[s];
            }
        }
    };
    u32x4 st[UPT];
    load_plane(i0 - 1, st);
    store_plane(0, st);
    load_plane(i0, st);
    store_plane(1, st);
    load_plane(i0 + 1, st);
    store_plane(2, st);
    __syncthreads();

    // apply epilogue: lane l serves voxel k0 + (l mod 32) of the wave's row and components 0-2
    // (l < 32) or 3-4; its inputs are loaded before the matrix work of the plane
    const int vx = lane & 31, half = lane >> 5, kx = k0 + vx;
    for (int i = i0; i < i1; ++i) {
        const bool more = i + 1 < i1;
        if (more) load_plane(i + 2, st);  // in flight during the matrix work of plane i
        bool alive = false;
        float qo[3] = {}, ro[3] = {};
        if constexpr (EPI == EPI_APPLY) {
            alive = j < Y && kx < Z && changeable(B, a.rows, i, j, kx);
            if (alive) {
                // the field is indexed by the box-local voxel (the activations by bbase + it)
                const long long vl = (long long)i * YZ + (long long)j * Z + kx;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int chn = half ? 3 + m : m;
                    if (chn < CH_Q) {
                        const long long idx = (long long)chn * a.cs + B.off + vl;
                        qo[m] = a.q_old[idx];
                        if (a.rhs) ro[m] = a.rhs[idx];
                    }
                }
            }
        }
        const int sl0 = (i - i0 + 1) % 3;  // slot of plane i (plane p sits in slot (p - i0 + 1) % 3)
        f32x4 acc[S][NB];
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[s][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            // tap and channel group of this lane's 8 K values
            const int t = CP == 32 ? ks : ks * 4 + g;
            const int gsel = CP == 32 ? g : 0;
            const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
            const int slot = (sl0 + 3 + di) % 3;
            const int base = slot * SLOT + ((w + 1 + dj) * G + gsel) * RW + col + 1 + dk;
            u32x4 wv[NB];
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) wv[nb] = wl[(nb * KS + ks) * 64 + lane];
#pragma unroll
            for (int s = 0; s < S; ++s) {
                u32x4 bv = {0u, 0u, 0u, 0u};
                if (t < TAPS) bv = ring[base + s * 16];
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) acc[s][nb] = mfma(wv[nb], bv, acc[s][nb]);
            }
        }

        lds_barrier();  // every wave is done with the slot of plane i-1
        if (more) store_plane((i + 3 - i0) % 3, st);  // plane i+2 -> slot (i+2 - i0 + 1) % 3
        // lane holds output channels nb*16 + g*4 + r (r = 0..3) of voxel (i, j, k0 + s*16 + col)
        if constexpr (EPI == EPI_APPLY) {
            // The five output channels of the wave's 16S voxels go through its LDS area, so that
            // lane l then serves voxel l mod 32 and components 0-2 (l < 32) or 3-4: every access
            // covers 32 consecutive voxels of one component plane (the values are unchanged).
            static_assert(S == 2, "the apply epilogue maps 32 voxels onto the two half-waves");
            float* yw = ytr + w * (CH_Q * 16 * S);
#pragma unroll
            for (int s = 0; s < S; ++s)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int chn = g * 4 + r;
                    if (chn < CH_Q) yw[chn * 16 * S + s * 16 + col] = acc[s][0][r] + a.bias[chn];
                }
            if (alive) {
                const long long vl = (long long)i * YZ + (long long)j * Z + kx;
                const bool face = i == 0 || j == 0 || kx == 0 || i == B.Nx - 1 || j == B.Ny - 1 || kx == B.Nz - 1;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int chn = half ? 3 + m : m;
                    if (chn < CH_Q) {
                        const long long idx = (long long)chn * a.cs + B.off + vl;
                        // one gain per voxel (u (1 + y_0) in every component) or one per component
                        const float y = yw[(a.scalar_gain ? 0 : chn) * 16 * S + vx];
                        float d = y;
                        if (a.rhs)  // increment input: the step's own increment, modulated
                            d = face ? 0.0f : (ro[m] - qo[m]) * (1.0f + y);
                        a.q_new[idx] = __builtin_fmaf(a.out_scale, d, qo[m]);
                    }
                }
            }
        } else if (j < Y) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const int k = k0 + s * 16 + col;
                if (k >= Z) continue;
                const long long v = bbase + (long long)i * YZ + (long long)j * Z + k;
#pragma unroll
                for (int nb = 0; nb < NB; ++nb) {
                    const int co = nb * 16 + g * 4;
                    const f32x4 c = acc[s][nb];
                    if constexpr (EPI == EPI_HIDDEN) {
                        const f32x4 bz = *reinterpret_cast<const f32x4*>(a.bias + co);
                        const f32x4 z = c + bz;
                        if (a.z_out) *reinterpret_cast<f32x4*>(a.z_out + v * 32 + co) = z;
                        uint16_t h[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) h[r] = bf16_bits(gelu_f(z[r]));
                        const uint2 pk = make_uint2(unsigned(h[0]) | (unsigned(h[1]) << 16),
                                                    unsigned(h[2]) | (unsigned(h[3]) << 16));
                        *reinterpret_cast<uint2*>(a.a_out + v * 32 + co) = pk;
                        if (a.a_cm) {
#pragma unroll
                            for (int r = 0; r < 4; ++r) a.a_cm[(long long)(co + r) * a.V + v] = h[r];
                        }
                    } else if constexpr (EPI == EPI_PRED) {
                        if (co < 8) {
                            const f32x4 bz = *reinterpret_cast<const f32x4*>(a.bias + co);
                            *reinterpret_cast<f32x4*>(a.pred + v * 8 + co) = c + bz;
                        }
                    } else if constexpr (EPI == EPI_BWD) {
                        const f32x4 z = *reinterpret_cast<const f32x4*>(a.z_in + v * 32 + co);
                        uint16_t h[4];
#pragma unroll
                        for (int r = 0; r < 4; ++r) h[r] = bf16_bits(c[r] * gelu_d(z[r]));
                        const uint2 pk = make_uint2(unsigned(h[0]) | (unsigned(h[1]) << 16),
                                                    unsigned(h[2]) | (unsigned(h[3]) << 16));
                        *reinterpret_cast<uint2*>(a.a_out + v * 32 + co) = pk;
#pragma unroll
                        for (int r = 0; r < 4; ++r) a.a_cm[(long long)(co + r) * a.V + v] = h[r];
                    }
                }
            }
        }
        lds_barrier();  // plane i+2 is in place before the next plane's reads
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[t][ci][co] = sum_v act[v + off(t), ci] * dy[v, co] and
// db[co] = sum_v dy[v, co], as a GEMM with M = (tap, ci) rows (plus one row of
// ones for the bias), N = output channels, K = voxels in chunks of 32.  Lane l
// holds 8 consecutive voxels (along k, Z % 8 == 0) of one row or channel:
// both operands are read from the channel-major copies.  A tap's k shift of
// +-1 element is rebuilt from aligned loads with v_alignbyte.  Each block
// accumulates over a grid-stride set of chunks and writes its partial sums.
// ---------------------------------------------------------------------------
__device__ __forceinline__ u32x4 load_run(const uint16_t* rowp, int k0, int dk, int Z) {
    const u32x4 cur = *reinterpret_cast<const u32x4*>(rowp + k0);
    if (dk == 0) return cur;
    u32x4 o;
    if (dk < 0) {
        const unsigned prev = k0 > 0 ? *reinterpret_cast<const unsigned*>(rowp + k0 - 2) : 0u;
        o[0] = __builtin_amdgcn_alignbyte(cur[0], prev, 2);
        o[1] = __builtin_amdgcn_alignbyte(cur[1], cur[0], 2);
        o[2] = __builtin_amdgcn_alignbyte(cur[2], cur[1], 2);
        o[3] = __builtin_amdgcn_alignbyte(cur[3], cur[2], 2);
    } else {
        const unsigned next = k0 + 8 < Z ? *reinterpret_cast<const unsigned*>(rowp + k0 + 8) : 0u;
        o[0] = __builtin_amdgcn_alignbyte(cur[1], cur[0], 2);
        o[1] = __builtin_amdgcn_alignbyte(cur[2], cur[1], 2);
        o[2] = __builtin_amdgcn_alignbyte(cur[3], cur[2], 2);
        o[3] = __builtin_amdgcn_alignbyte(next, cur[3], 2);
    }
    return o;
}

template <int CP, int NB>
__global__ __launch_bounds__(256) void wgrad_kernel(const uint16_t* __restrict__ act_cm,
                                                    const uint16_t* __restrict__ dy_cm, int X, int Y, int Z,
                                                    long long V, float* __restrict__ part) {
    constexpr int ROWS = TAPS * CP;
    constexpr int MBT = (ROWS + 1 + 15) / 16;
    constexpr int MBW = (MBT + 3) / 4;
    constexpr int COP = NB == 1 ? 8 : 32;  // output-gradient channels stored
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, col = lane & 15;
    f32x4 acc[MBW][NB];
#pragma unroll
    for (int m = 0; m < MBW; ++m)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[m][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    const long long nchunk = (V + 31) / 32;
    for (long long chk = blockIdx.x; chk < nchunk; chk += gridDim.x) {
        const long long v0 = chk * 32 + g * 8;
        const bool run_ok = v0 < V;
        const long long row = run_ok ? v0 / Z : 0;
        const int k0 = int(v0 - row * Z);
        const int j = int(row % Y), i = int((row / Y) % X);
        u32x4 bfr[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const int co = nb * 16 + col;
            bfr[nb] = u32x4{0u, 0u, 0u, 0u};
            if (run_ok && co < COP) bfr[nb] = *reinterpret_cast<const u32x4*>(dy_cm + (long long)co * V + v0);
        }
#pragma unroll
        for (int m = 0; m < MBW; ++m) {
            const int mb = wave + 4 * m;
            if (mb >= MBT) break;
            const int R = mb * 16 + col;
            u32x4 afr = {0u, 0u, 0u, 0u};
            if (run_ok) {
                if (R < ROWS) {
                    const int t = R / CP, ci = R % CP;
                    const int di = t / 9 - 1, dj = (t / 3) % 3 - 1, dk = t % 3 - 1;
                    if (unsigned(i + di) < unsigned(X) && unsigned(j + dj) < unsigned(Y)) {
                        const long long nrow = row + (long long)di * Y + dj;
                        afr = load_run(act_cm + (long long)ci * V + nrow * Z, k0, dk, Z);
                    }
                } else if (R == ROWS) {
                    afr = u32x4{0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u};  // BF16 1.0
                }
            }
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[m][nb] = mfma(afr, bfr[nb], acc[m][nb]);
        }
    }
    // lane holds rows mb*16 + g*4 + r, output channel nb*16 + col
#pragma unroll
    for (int m = 0; m < MBW; ++m) {
        const int mb = wave + 4 * m;
        if (mb >= MBT) break;
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const long long R = mb * 16 + g * 4 + r;
                part[((long long)blockIdx.x * MBT * 16 + R) * (NB * 16) + nb * 16 + col] = acc[m][nb][r];
            }
    }
}

// Fixed-order sum of the per-block partials into the flat gradient.
__global__ __launch_bounds__(256) void wgrad_reduce_kernel(const float* __restrict__ part, int nblk, int cp, int mbt,
                                                           int ncol, int cin, int cout, long long w_off,
                                                           long long b_off, float* __restrict__ grad) {
    const int rows = TAPS * cp;
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)(rows + 1) * cout;
    if (idx >= total) return;
    const int R = int(idx / cout), co = int(idx % cout);
    long long dst;
    if (R < rows) {
        const int t = R / cp, ci = R % cp;
        if (ci >= cin) return;
        dst = w_off + ((long long)t * cin + ci) * cout + co;
    } else {
        dst = b_off + co;
    }
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += part[((long long)b * mbt * 16 + R) * ncol + co];
    grad[dst] = float(s);
}

// ---------------------------------------------------------------------------
// Parameter packing: FP32 master parameters -> BF16 fragments and padded biases.
// Forward fragment (cp, nb): value W[t][ci][co], K = t*cp + ci, co = nb*16 + l%16.
// Input-gradient fragment of a layer (cp = its padded output count): the
// convolution of dy with W'[t][co][ci] = W[26 - t][ci][co].
// ---------------------------------------------------------------------------
struct FragSpec {
    uint16_t* dst;
    long long w_off;
    int cp, nb, cin, cout, transpose;
};

__device__ float frag_value(const float* P, const FragSpec& s, long long e) {
    const int KS = ksteps(s.cp);
    const int el = int(e & 7), l = int((e >> 3) & 63);
    const long long fi = e >> 9;  // nb * KS + ks
    const int ks = int(fi % KS), nb = int(fi / KS);
    const int K = ks * 32 + (l >> 4) * 8 + el;
    const int t = K / s.cp, kc = K % s.cp;
    const int oc = nb * 16 + (l & 15);
    if (t >= TAPS) return 0.0f;
    if (!s.transpose) {
        if (kc >= s.cin || oc >= s.cout) return 0.0f;
        return P[s.w_off + ((long long)t * s.cin + kc) * s.cout + oc];
    }
    // input gradient: kc runs over the layer's outputs, oc over its inputs
    if (kc >= s.cout || oc >= s.cin) return 0.0f;
    return P[s.w_off + ((long long)(TAPS - 1 - t) * s.cin + oc) * s.cout + kc];
}

__global__ __launch_bounds__(256) void pack_frags_kernel(const float* __restrict__ P, Frags f) {
    const FragSpec spec[5] = {
        {f.f1, OFF_W1, 8, 2, CH_Q, CH_H, 0},   {f.f2, OFF_W2, 32, 2, CH_H, CH_H, 0},
        {f.f3, OFF_W3, 32, 1, CH_H, CH_Q, 0},  {f.g3, OFF_W3, 8, 2, CH_H, CH_Q, 1},
        {f.g2, OFF_W2, 32, 2, CH_H, CH_H, 1},
    };
    long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    for (int s = 0; s < 5; ++s) {
        const long long n = frag_elems(spec[s].cp, spec[s].nb);
        if (e < n) {
            spec[s].dst[e] = bf16_bits(frag_value(P, spec[s], e));
            return;
        }
        e -= n;
    }
    if (e < 32) {
        f.b1[e] = P[OFF_B1 + e];
        return;
    }
    e -= 32;
    if (e < 32) {
        f.b2[e] = P[OFF_B2 + e];
        return;
    }
    e -= 32;
    if (e < 16) f.b3[e] = e < CH_Q ? P[OFF_B3 + e] : 0.0f;
}

// Adam (Kingma & Ba 2015), bias corrections folded into lr_t and inv_sqrt_bc2:
// p -= lr_t * m / (sqrt(v) * inv_sqrt_bc2 + eps),  lr_t = lr / (1 - b1^t),
// inv_sqrt_bc2 = 1 / sqrt(1 - b2^t).
__global__ __launch_bounds__(256) void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, long long n,
                                                   float lr_t, float b1, float b2, float eps, float inv_sqrt_bc2) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) * inv_sqrt_bc2 + eps);
}

// ---------------------------------------------------------------------------
// Batch assembly
// ---------------------------------------------------------------------------
__device__ __forceinline__ void store_x(const Work& w, long long v, const float q[5]) {
    uint16_t h[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) h[c] = c < CH_Q ? bf16_bits(q[c]) : uint16_t(0);
    u32x4 pk;
#pragma unroll
    for (int c = 0; c < 4; ++c) pk[c] = unsigned(h[2 * c]) | (unsigned(h[2 * c + 1]) << 16);
    *reinterpret_cast<u32x4*>(w.x + v * 8) = pk;
#pragma unroll
    for (int c = 0; c < 8; ++c) w.x_cm[(long long)c * w.V + v] = h[c];
}

__global__ __launch_bounds__(256) void crop_kernel(const Snap* __restrict__ snaps, const CropJob* __restrict__ jobs,
                                                   int C, Work w) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= w.V) return;
    const long long C3 = (long long)C * C * C;
    const int b = int(v / C3);
    const long long loc = v - b * C3;
    const int i = int(loc / ((long long)C * C)), j = int((loc / C) % C), k = int(loc % C);
    const CropJob J = jobs[b];
    const Snap si = snaps[J.s_in], so = snaps[J.s_out];
    const long long src = ((long long)(J.ox + i) * si.Y + (J.oy + j)) * si.Z + (J.oz + k);
    float xi[5], mi[5], d[8];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const float qi = si.q[c * si.nvox + src];
        d[c] = (so.q[c * so.nvox + src] - qi) * J.out_scale;
        mi[c] = J.kind ? si.inc[c * si.nvox + src] * J.in_scale : 0.0f;
        xi[c] = (J.kind == 1 || J.kind == 3) ? mi[c] : qi;  // increment or step-size input read u / s
    }
    d[5] = d[6] = d[7] = 0.0f;
    store_x(w, v, xi);
    *reinterpret_cast<f32x4*>(w.mod + v * 8) = f32x4{mi[0], mi[1], mi[2], mi[3]};
    *reinterpret_cast<f32x4*>(w.mod + v * 8 + 4) = f32x4{mi[4], 0.0f, 0.0f, 0.0f};
    *reinterpret_cast<f32x4*>(w.tgt + v * 8) = f32x4{d[0], d[1], d[2], d[3]};
    *reinterpret_cast<f32x4*>(w.tgt + v * 8 + 4) = f32x4{d[4], d[5], d[6], d[7]};
    // a crop face that is not a box face sees zero padding instead of the true
    // neighbours: the 3-voxel band (the network's reach) behind it is not scored
    const int m3 = 3;
    const bool inner = (i >= m3 || J.ox == 0) && (i < C - m3 || J.ox + C == si.X) && (j >= m3 || J.oy == 0) &&
                       (j < C - m3 || J.oy + C == si.Y) && (k >= m3 || J.oz == 0) && (k < C - m3 || J.oz + C == si.Z);
    w.mask[v] = (inner && si.mask[src]) ? J.weight : 0.0f;
}

__global__ __launch_bounds__(256) void pack_batch_kernel(const float* __restrict__ x5, const float* __restrict__ t5,
                                                         const float* __restrict__ mk, const float* __restrict__ m5,
                                                         Work w) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= w.V) return;
    float q[5], d[8] = {};
#pragma unroll
    for (int c = 0; c < 5; ++c) q[c] = x5[v * 5 + c];
    store_x(w, v, q);
    float mo[5] = {};
    if (m5) {
#pragma unroll
        for (int c = 0; c < 5; ++c) mo[c] = m5[v * 5 + c];
    }
    *reinterpret_cast<f32x4*>(w.mod + v * 8) = f32x4{mo[0], mo[1], mo[2], mo[3]};
    *reinterpret_cast<f32x4*>(w.mod + v * 8 + 4) = f32x4{mo[4], 0.0f, 0.0f, 0.0f};
    if (t5) {
#pragma unroll
        for (int c = 0; c < 5; ++c) d[c] = t5[v * 5 + c];
    }
    *reinterpret_cast<f32x4*>(w.tgt + v * 8) = f32x4{d[0], d[1], d[2], d[3]};
    *reinterpret_cast<f32x4*>(w.tgt + v * 8 + 4) = f32x4{d[4], d[5], d[6], d[7]};
    w.mask[v] = mk ? mk[v] : 1.0f;
}

__global__ __launch_bounds__(256) void pack_field_kernel(const float* __restrict__ q, long long cs, long long off,
                                                         long long V, uint16_t* __restrict__ x) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    uint16_t h[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) h[c] = c < CH_Q ? bf16_bits(q[(long long)c * cs + off + v]) : uint16_t(0);
    u32x4 pk;
#pragma unroll
    for (int c = 0; c < 4; ++c) pk[c] = unsigned(h[2 * c]) | (unsigned(h[2 * c + 1]) << 16);
    *reinterpret_cast<u32x4*>(x + v * 8) = pk;
}

__global__ __launch_bounds__(256) void unpack_pred_kernel(Work w, float* __restrict__ out5) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= w.V) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const float y = w.pred[v * 8 + (w.gain == 2 ? 0 : c)];
        out5[v * 5 + c] = w.gain ? w.mod[v * 8 + c] * (1.0f + y) : y;
    }
}

// ---------------------------------------------------------------------------
// Reductions (FP64, fixed order: per-block partials, then one finishing block)
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void block_sum_store(double (&v)[N], double* dst) {
    __shared__ double red[N][256];
#pragma unroll
    for (int s = 0; s < N; ++s) red[s][threadIdx.x] = v[s];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
#pragma unroll
            for (int s = 0; s < N; ++s) red[s][threadIdx.x] += red[s][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int s = 0; s < N; ++s) dst[s] = red[s][0];
    }
}

__global__ __launch_bounds__(256) void loss_kernel(Work w, float scale) {
    double acc[2] = {0.0, 0.0};
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < w.V;
         v += (long long)gridDim.x * blockDim.x) {
        const float mk = w.mask[v];
        uint16_t h[8];
        if (w.gain == 2) {  // prediction m (1 + y_0) in every component: dL/dy_0 = sum_c dL/dp_c m_c
            const float y0 = w.pred[v * 8];
            float g0 = 0.0f;
#pragma unroll
            for (int c = 0; c < CH_Q; ++c) {
                const float m = w.mod[v * 8 + c];
                const float d = m * (1.0f + y0) - w.tgt[v * 8 + c];
                acc[0] += double(mk) * double(d) * double(d);
                g0 = __builtin_fmaf(d, m, g0);
            }
            h[0] = bf16_bits(scale * mk * g0);
#pragma unroll
            for (int c = 1; c < 8; ++c) h[c] = 0;
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                float d = 0.0f, g = 1.0f;
                if (c < CH_Q) {
                    const float y = w.pred[v * 8 + c];
                    if (w.gain) {  // prediction m (1 + y): dL/dy = dL/dp * m
                        g = w.mod[v * 8 + c];
                        d = g * (1.0f + y) - w.tgt[v * 8 + c];
                    } else {
                        d = y - w.tgt[v * 8 + c];
                    }
                }
                acc[0] += double(mk) * double(d) * double(d);
                h[c] = bf16_bits(scale * mk * d * g);
            }
        }
        acc[1] += mk != 0.0f ? 1.0 : 0.0;  // scored voxels (mask values may be weights)
        u32x4 pk;
#pragma unroll
        for (int c = 0; c < 4; ++c) pk[c] = unsigned(h[2 * c]) | (unsigned(h[2 * c + 1]) << 16);
        *reinterpret_cast<u32x4*>(w.dy + v * 8) = pk;
#pragma unroll
        for (int c = 0; c < 8; ++c) w.dy_cm[(long long)c * w.V + v] = h[c];
    }
    block_sum_store<2>(acc, w.red + 2 * blockIdx.x);
}

// Fixed-order sum of nblk x nslot partials into out[nslot], by one block.
__device__ __forceinline__ void sum_slots_block(const double* __restrict__ partials, long long nblk, int nslot,
                                                double* __restrict__ out) {
    __shared__ double red[256];
    for (int s = 0; s < nslot; ++s) {
        double a = 0.0;
        for (long long b = threadIdx.x; b < nblk; b += 256) a += partials[b * nslot + s];
        red[threadIdx.x] = a;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[s] = red[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void sum_slots_kernel(const double* __restrict__ partials, long long nblk,
                                                        int nslot, double* __restrict__ out) {
    sum_slots_block(partials, nblk, nslot, out);
}

// The same for n lists: block m sums partials [prefix[m], prefix[m+1]) into out + m * nslot.
__global__ __launch_bounds__(256) void sum_slots_multi_kernel(const double* __restrict__ partials,
                                                              const long long* __restrict__ prefix, int nslot,
                                                              double* __restrict__ out) {
    const int m = blockIdx.x;
    sum_slots_block(partials + prefix[m] * nslot, prefix[m + 1] - prefix[m], nslot, out + (long long)m * nslot);
}

// Global block g of a multi-box launch over prefix[0..n] (prefix[m] = first block of box m):
// its box m and its block index within the box.
__device__ __forceinline__ int block_box(long long g, const long long* __restrict__ prefix, int n) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(256) void pair_power_kernel(const float* __restrict__ qa, const float* __restrict__ qb,
                                                         const uint8_t* __restrict__ mask, long long nvox,
                                                         double* __restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long long)gridDim.x * blockDim.x) {
        if (!mask[v]) continue;
#pragma unroll
        for (int c = 0; c < CH_Q; ++c) {
            const double d = double(qb[c * nvox + v]) - double(qa[c * nvox + v]);
            acc[0] += d * d;
        }
        acc[1] += 1.0;
    }
    block_sum_store<2>(acc, partials + 2 * blockIdx.x);
}

__global__ __launch_bounds__(256) void mask_bytes_kernel(long long nvox, BoxDev B, const RowMask* __restrict__ rows,
                                                         uint8_t* __restrict__ out) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int k = int(v % B.Nz), j = int((v / B.Nz) % B.Ny), i = int(v / ((long long)B.Nz * B.Ny));
    out[v] = changeable(B, rows, i, j, k) ? 1 : 0;
}

// One block's share (block bid of nblk, grid stride over the box's voxels) of the residual
// partials: sum of |rhs - q|^2 and the voxel count.
__device__ __forceinline__ void residual_block(const float* __restrict__ q, const float* __restrict__ rhs,
                                               long long cs, const BoxDev& B, const RowMask* __restrict__ rows,
                                               double* __restrict__ partials, long long bid, long long nblk) {
    const long long nvox = (long long)B.Nx * B.Ny * B.Nz;
    double acc[2] = {0.0, 0.0};
    for (long long v = bid * blockDim.x + threadIdx.x; v < nvox; v += nblk * blockDim.x) {
        const int k = int(v % B.Nz), j = int((v / B.Nz) % B.Ny), i = int(v / ((long long)B.Nz * B.Ny));
        if (!stepped_interior(B, rows, i, j, k)) continue;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const long long idx = (long long)c * cs + B.off + v;
            const double d = double(rhs[idx]) - double(q[idx]);
            acc[0] += d * d;
        }
        acc[1] += 1.0;
    }
    block_sum_store<2>(acc, partials + 2 * bid);
}

__global__ __launch_bounds__(256) void residual_kernel(const float* __restrict__ q, const float* __restrict__ rhs,
                                                       long long cs, BoxDev B, const RowMask* __restrict__ rows,
                                                       double* __restrict__ partials) {
    residual_block(q, rhs, cs, B, rows, partials, blockIdx.x, gridDim.x);
}

// Residual partials of n boxes in one launch: global block g belongs to box m = boxes[...] with
// prefix[m] <= g < prefix[m+1], and computes exactly block g - prefix[m] of residual_kernel's
// grid for that box; its partials go to partials + 2 g.
__global__ __launch_bounds__(256) void residual_multi_kernel(const float* __restrict__ q,
                                                             const float* __restrict__ rhs, long long cs,
                                                             const BoxDev* __restrict__ boxes,
                                                             const int* __restrict__ list,
                                                             const long long* __restrict__ prefix, int n,
                                                             const RowMask* __restrict__ rows,
                                                             double* __restrict__ partials) {
    const long long g = blockIdx.x;
    const int m = block_box(g, prefix, n);
    const BoxDev B = boxes[list[m]];
    residual_block(q, rhs, cs, B, rows, partials + 2 * prefix[m], g - prefix[m], prefix[m + 1] - prefix[m]);
}

// The increment of one explicit step, rhs - q, on the voxels the residual sums
// over (zero elsewhere: rhs is not written on static voxels), per box voxel:
// five planes of nvox floats (a dataset snapshot's layout).
__global__ __launch_bounds__(256) void increment_box_kernel(const float* __restrict__ q, const float* __restrict__ rhs,
                                                            long long cs, BoxDev B, const RowMask* __restrict__ rows,
                                                            float* __restrict__ dst) {
    const long long nvox = (long long)B.Nx * B.Ny * B.Nz;
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int k = int(v % B.Nz), j = int((v / B.Nz) % B.Ny), i = int(v / ((long long)B.Nz * B.Ny));
    const bool on = stepped_interior(B, rows, i, j, k);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const long long idx = (long long)c * cs + B.off + v;
        dst[(long long)c * nvox + v] = on ? rhs[idx] - q[idx] : 0.0f;
    }
}

// Network input from the increment: x = bf16((rhs - q) * inv_s) (5 of 8
// channels) on the voxels the residual sums over, zero elsewhere.
__global__ __launch_bounds__(256) void pack_increment_kernel(const float* __restrict__ q, const float* __restrict__ rhs,
                                                             long long cs, BoxDev B, const RowMask* __restrict__ rows,
                                                             float inv_s, uint16_t* __restrict__ x) {
    const long long nvox = (long long)B.Nx * B.Ny * B.Nz;
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int k = int(v % B.Nz), j = int((v / B.Nz) % B.Ny), i = int(v / ((long long)B.Nz * B.Ny));
    const bool on = stepped_interior(B, rows, i, j, k);
    uint16_t h[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const long long idx = (long long)c * cs + B.off + v;
        h[c] = (c < CH_Q && on) ? bf16_bits((rhs[idx] - q[idx]) * inv_s) : uint16_t(0);
    }
    u32x4 pk;
#pragma unroll
    for (int c = 0; c < 4; ++c) pk[c] = unsigned(h[2 * c]) | (unsigned(h[2 * c + 1]) << 16);
    *reinterpret_cast<u32x4*>(x + v * 8) = pk;
}

// A proposed step taken back to a fraction of itself: other = cur + ratio (other - cur)
// on the voxels a step can change (where the proposal wrote).
__global__ __launch_bounds__(256) void rescale_step_kernel(const float* __restrict__ cur, float* __restrict__ other,
                                                           long long cs, BoxDev B, const RowMask* __restrict__ rows,
                                                           float ratio) {
    const long long nvox = (long long)B.Nx * B.Ny * B.Nz;
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int k = int(v % B.Nz), j = int((v / B.Nz) % B.Ny), i = int(v / ((long long)B.Nz * B.Ny));
    if (!changeable(B, rows, i, j, k)) return;
#pragma unroll
    for (int c = 0; c < CH_Q; ++c) {
        const long long idx = (long long)c * cs + B.off + v;
        other[idx] = __builtin_fmaf(ratio, other[idx] - cur[idx], cur[idx]);
    }
}

__global__ __launch_bounds__(256) void copy_box_kernel(const float* __restrict__ q, long long cs, long long off,
                                                       long long nvox, float* __restrict__ dst) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) dst[(long long)c * nvox + v] = q[(long long)c * cs + off + v];
}

unsigned blocks_for(long long n) { return unsigned((n + 255) / 256); }
unsigned reduce_blocks(long long n) { return std::max(1u, std::min(blocks_for(n), 2048u)); }

template <int CP, int NB, int EPI, int S, int TJ>
hipError_t go_conv_s(hipStream_t st, const ConvArgs& a) {
    using T = ConvTile<CP, S, TJ>;
    const size_t lds = size_t(conv_lds_units<CP, NB, EPI, S, TJ>()) * 16;
    int ntj = 0, ntk = 0, ch = 0, nti = 0;
    long long grid = 0;
    if (a.tiles) {
        // a tile list: one block per listed tile
        if (T::TJ != TILE_TJ || T::TK != TILE_TK || a.tile_ch < 1) return hipErrorInvalidValue;
        ch = a.tile_ch;
        grid = a.ntiles;
    } else {
        ntj = (a.Y + T::TJ - 1) / T::TJ;
        ntk = (a.Z + T::TK - 1) / T::TK;
        // planes per block: aim at about four blocks per CU over the whole launch
        const long long tiles = (long long)a.n * ntj * ntk;
        const long long want_nti = std::max(1LL, (1024 + tiles - 1) / tiles);
        ch = int(std::max(2LL, (a.X + want_nti - 1) / want_nti));
        nti = (a.X + ch - 1) / ch;
        grid = tiles * nti;
    }
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
    // allow the kernel more than the default dynamic LDS; a device that cannot
    // give it makes the launch below fail with an error
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&conv_kernel<CP, NB, EPI, S, TJ>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    hipLaunchKernelGGL((conv_kernel<CP, NB, EPI, S, TJ>), dim3(unsigned(grid)), dim3(T::NT), lds, st, a, ch, ntj,
                       ntk, nti);
    return hipGetLastError();
}

template <int CP, int NB, int EPI>
hipError_t go_conv(hipStream_t st, const ConvArgs& a) {
    // 32 voxels of k per block and one wave per row: 12 rows (768 threads, three waves per SIMD,
    // 146 KB of LDS for 32 -> 32) on large boxes, 8 rows on short ones (training crops)
    // (tile lists use the 12-row shape)
    return a.tiles || a.Y > 32 ? go_conv_s<CP, NB, EPI, 2, TILE_TJ>(st, a) : go_conv_s<CP, NB, EPI, 2, 8>(st, a);
}

}  // namespace

namespace qsrs {

hipError_t launch_conv(int cp, int nb, int epi, hipStream_t st, const ConvArgs& a) {
    if (cp == 8 && nb == 2 && epi == EPI_HIDDEN) return go_conv<8, 2, EPI_HIDDEN>(st, a);
    if (cp == 32 && nb == 2 && epi == EPI_HIDDEN) return go_conv<32, 2, EPI_HIDDEN>(st, a);
    if (cp == 32 && nb == 1 && epi == EPI_PRED) return go_conv<32, 1, EPI_PRED>(st, a);
    if (cp == 32 && nb == 1 && epi == EPI_APPLY) return go_conv<32, 1, EPI_APPLY>(st, a);
    if (cp == 8 && nb == 2 && epi == EPI_BWD) return go_conv<8, 2, EPI_BWD>(st, a);
    if (cp == 32 && nb == 2 && epi == EPI_BWD) return go_conv<32, 2, EPI_BWD>(st, a);
    return hipErrorInvalidValue;
}

hipError_t launch_wgrad(int cp, int nb, hipStream_t st, const uint16_t* act_cm, const uint16_t* dz_cm, int X, int Y,
                        int Z, long long V, float* part) {
    if (Z % 8 != 0) return hipErrorInvalidValue;
    if (cp == 8 && nb == 2)
        hipLaunchKernelGGL((wgrad_kernel<8, 2>), dim3(WG_BLOCKS), dim3(256), 0, st, act_cm, dz_cm, X, Y, Z, V, part);
    else if (cp == 32 && nb == 2)
        hipLaunchKernelGGL((wgrad_kernel<32, 2>), dim3(WG_BLOCKS), dim3(256), 0, st, act_cm, dz_cm, X, Y, Z, V, part);
    else if (cp == 32 && nb == 1)
        hipLaunchKernelGGL((wgrad_kernel<32, 1>), dim3(WG_BLOCKS), dim3(256), 0, st, act_cm, dz_cm, X, Y, Z, V, part);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_wgrad_reduce(int cp, int nb, int cin_real, int cout_real, long long w_off, long long b_off,
                               hipStream_t st, const float* part, float* grad) {
    const int mbt = (TAPS * cp + 1 + 15) / 16;
    const long long total = (long long)(TAPS * cp + 1) * cout_real;
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(blocks_for(total)), dim3(256), 0, st, part, WG_BLOCKS, cp, mbt,
                       nb * 16, cin_real, cout_real, w_off, b_off, grad);
    return hipGetLastError();
}

hipError_t launch_pack_frags(hipStream_t st, const float* params, Frags f) {
    const long long n = frag_elems(8, 2) + frag_elems(32, 2) + frag_elems(32, 1) + frag_elems(8, 2) +
                        frag_elems(32, 2) + 32 + 32 + 16;
    hipLaunchKernelGGL(pack_frags_kernel, dim3(blocks_for(n)), dim3(256), 0, st, params, f);
    return hipGetLastError();
}

hipError_t launch_adam(hipStream_t st, float* params, const float* grad, float* m, float* v, long long n, float lr_t,
                       float b1, float b2, float eps, float inv_sqrt_bc2) {
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(n)), dim3(256), 0, st, params, grad, m, v, n, lr_t, b1, b2, eps,
                       inv_sqrt_bc2);
    return hipGetLastError();
}

hipError_t launch_crop(hipStream_t st, const Snap* snaps, const CropJob* jobs, int nbatch, int C, const Work& w) {
    (void)nbatch;
    hipLaunchKernelGGL(crop_kernel, dim3(blocks_for(w.V)), dim3(256), 0, st, snaps, jobs, C, w);
    return hipGetLastError();
}

hipError_t launch_pack_batch(hipStream_t st, const float* x5, const float* t5, const float* m, const float* m5,
                             long long V, const Work& w) {
    (void)V;
    hipLaunchKernelGGL(pack_batch_kernel, dim3(blocks_for(w.V)), dim3(256), 0, st, x5, t5, m, m5, w);
    return hipGetLastError();
}

hipError_t launch_pack_field(hipStream_t st, const float* q, long long cs, long long off, long long V, uint16_t* x) {
    hipLaunchKernelGGL(pack_field_kernel, dim3(blocks_for(V)), dim3(256), 0, st, q, cs, off, V, x);
    return hipGetLastError();
}

hipError_t launch_unpack_pred(hipStream_t st, const Work& w, float* out5) {
    hipLaunchKernelGGL(unpack_pred_kernel, dim3(blocks_for(w.V)), dim3(256), 0, st, w, out5);
    return hipGetLastError();
}

hipError_t launch_loss(hipStream_t st, const Work& w, float scale, long long* nblk) {
    const unsigned g = reduce_blocks(w.V);
    *nblk = g;
    hipLaunchKernelGGL(loss_kernel, dim3(g), dim3(256), 0, st, w, scale);
    return hipGetLastError();
}

hipError_t launch_sum_slots(hipStream_t st, const double* partials, long long nblk, int nslot, double* out) {
    hipLaunchKernelGGL(sum_slots_kernel, dim3(1), dim3(256), 0, st, partials, nblk, nslot, out);
    return hipGetLastError();
}

hipError_t launch_pair_power(hipStream_t st, const float* qa, const float* qb, const uint8_t* mask, long long nvox,
                             double* partials, long long* nblk) {
    const unsigned g = reduce_blocks(nvox);
    *nblk = g;
    hipLaunchKernelGGL(pair_power_kernel, dim3(g), dim3(256), 0, st, qa, qb, mask, nvox, partials);
    return hipGetLastError();
}

hipError_t launch_mask_bytes(hipStream_t st, long long nvox, const qsrk::BoxDev& box, const qsrk::RowMask* rows,
                             uint8_t* out) {
    hipLaunchKernelGGL(mask_bytes_kernel, dim3(blocks_for(nvox)), dim3(256), 0, st, nvox, box, rows, out);
    return hipGetLastError();
}

hipError_t launch_residual(hipStream_t st, const float* q, const float* rhs, long long cs, const qsrk::BoxDev& box,
                           const qsrk::RowMask* rows, double* partials, long long* nblk) {
    const long long nvox = (long long)box.Nx * box.Ny * box.Nz;
    const unsigned g = reduce_blocks(nvox);
    *nblk = g;
    hipLaunchKernelGGL(residual_kernel, dim3(g), dim3(256), 0, st, q, rhs, cs, box, rows, partials);
    return hipGetLastError();
}

long long residual_blocks(long long nvox) { return reduce_blocks(nvox); }

hipError_t launch_residual_multi(hipStream_t st, const float* q, const float* rhs, long long cs,
                                 const qsrk::BoxDev* boxes, const int* list, const long long* prefix, int n,
                                 long long total_blocks, const qsrk::RowMask* rows, double* partials, double* out) {
    if (n <= 0 || total_blocks <= 0) return hipSuccess;
    if (total_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(residual_multi_kernel, dim3(unsigned(total_blocks)), dim3(256), 0, st, q, rhs, cs, boxes, list,
                       prefix, n, rows, partials);
    hipLaunchKernelGGL(sum_slots_multi_kernel, dim3(unsigned(n)), dim3(256), 0, st, partials, prefix, 2, out);
    return hipGetLastError();
}

hipError_t launch_increment_box(hipStream_t st, const float* q, const float* rhs, long long cs,
                                const qsrk::BoxDev& box, const qsrk::RowMask* rows, float* dst) {
    const long long nvox = (long long)box.Nx * box.Ny * box.Nz;
    hipLaunchKernelGGL(increment_box_kernel, dim3(blocks_for(nvox)), dim3(256), 0, st, q, rhs, cs, box, rows, dst);
    return hipGetLastError();
}

hipError_t launch_pack_increment(hipStream_t st, const float* q, const float* rhs, long long cs,
                                 const qsrk::BoxDev& box, const qsrk::RowMask* rows, float inv_s, uint16_t* x) {
    const long long nvox = (long long)box.Nx * box.Ny * box.Nz;
    hipLaunchKernelGGL(pack_increment_kernel, dim3(blocks_for(nvox)), dim3(256), 0, st, q, rhs, cs, box, rows, inv_s,
                       x);
    return hipGetLastError();
}

hipError_t launch_rescale_step(hipStream_t st, const float* cur, float* other, long long cs, const qsrk::BoxDev& box,
                               const qsrk::RowMask* rows, float ratio) {
    const long long nvox = (long long)box.Nx * box.Ny * box.Nz;
    hipLaunchKernelGGL(rescale_step_kernel, dim3(blocks_for(nvox)), dim3(256), 0, st, cur, other, cs, box, rows,
                       ratio);
    return hipGetLastError();
}

hipError_t launch_copy_box(hipStream_t st, const float* q, long long cs, long long off, long long nvox, float* dst) {
    hipLaunchKernelGGL(copy_box_kernel, dim3(blocks_for(nvox)), dim3(256), 0, st, q, cs, off, nvox, dst);
    return hipGetLastError();
}

}  // namespace qsrs
