// qsr_step.hip -- fused FP32 LdG step kernels for gfx950.
//
// One launch advances every box by one explicit-Euler iteration:
//   dQ = -(dt/gamma) * P[ h_bulk - L1 lap Q - L2 sym d(div Q) + L3 terms ]
//        (+ W/delta (Q - Q0) on the shell) (+ noise), strong shell re-imposed.
// This fuses the reference's per-iteration kernel chain (QSR.cu:279-867,
// launched at QSR.cu:3223-3238) and the periodic bulk kernel
// (KZM_bulk_ldg.cu:351-469) into one pass that reads Q once and writes Q once.
//
// Layout: 5 FP32 SoA planes (Qxx, Qxy, Qxz, Qyy, Qyz), each `cs` floats, box b
// at offset boxes[b].off, voxel (i,j,k) at k + Nz*(j + Ny*i).  Ping-pong
// buffers; voxels that can never change are not written (both buffers hold
// them from upload).
//
// Tiles: a block owns TJ rows x TK consecutive k (64 x 4 or 32 x 8) of
// `ch` consecutive i-planes and streams along i.  Box faces of clamped boxes
// are excluded here and done by step_face_kernel (one-sided differences).
#include "qsr_kernels.hpp"

using namespace qsrk;

namespace {

__device__ __forceinline__ int wrap_or_clamp(int x, int n, bool periodic) {
    if (periodic) {
        x %= n;
        return x < 0 ? x + n : x;
    }
    return x < 0 ? 0 : (x >= n ? n - 1 : x);
}

__device__ __forceinline__ bool in_iv(int k, short lo, short hi) { return k >= lo && k <= hi; }

// A value equal in every lane, moved to a scalar register.
__device__ __forceinline__ float uni(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

// Workgroup barrier that orders LDS only.  The step kernels never exchange data
// through global memory inside a launch (they read one buffer and write the
// other), so the barrier need not drain outstanding global loads: prefetches of
// the next planes stay in flight across it (a full __syncthreads() waits for
// vmcnt(0) first).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Field access for the stencils.  The context's 5 SoA planes are addressed as
// base(box) + component * cs + plane offset + in-plane offset.  BUF: raw buffer
// loads/stores from the box base, the byte offset split into a per-lane part
// (in-plane position, fixed for a thread) and a wave-uniform part (component and
// i-plane), so an access costs no vector address arithmetic; valid when every
// offset of the context fits 32 bits (20 B x cs < 4 GiB).  Otherwise flat
// 64-bit addressing.  `st` stores are non-temporal; `st_cached` stores are
// ordinary ones, which the memory-side cache keeps for the next launch's reads
// (the two-step kernel's: 0.156 ms instead of 0.172 ms per launch at 256^3,
// while the single-step kernels are faster with non-temporal stores).
template <bool BUF>
struct Field;

template <>
struct Field<false> {
    const float* in;
    float* out;
    long long cs;
    __device__ Field(const float* qin, float* qout, long long cs_, long long off)
        : in(qin + off), out(qout + off), cs(cs_) {}
    __device__ float ld(int c, long long pbase, long long o) const { return in[c * cs + pbase + o]; }
    __device__ void st(int c, long long pbase, long long o, float v) const {
        __builtin_nontemporal_store(v, out + c * cs + pbase + o);
    }
    __device__ void st_cached(int c, long long pbase, long long o, float v) const { out[c * cs + pbase + o] = v; }
};

template <>
struct Field<true> {
    static constexpr int kDword3 = 0x00020000;  // raw 32-bit access, gfx9 buffer descriptor
    __amdgpu_buffer_rsrc_t rin, rout;
    unsigned cs4;
    __device__ Field(const float* qin, float* qout, long long cs_, long long off)
        : rin(__builtin_amdgcn_make_buffer_rsrc((void*)(qin + off), 0, -1, kDword3)),
          rout(__builtin_amdgcn_make_buffer_rsrc((void*)(qout + off), 0, -1, kDword3)),
          cs4(unsigned(cs_) * 4u) {}
    __device__ float ld(int c, long long pbase, long long o) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rin, unsigned(o) * 4u,
                                                                              unsigned(c) * cs4 + unsigned(pbase) * 4u, 0));
    }
    __device__ void st(int c, long long pbase, long long o, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rout, unsigned(o) * 4u,
                                              unsigned(c) * cs4 + unsigned(pbase) * 4u, 2 /* nt */);
    }
    __device__ void st_cached(int c, long long pbase, long long o, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), rout, unsigned(o) * 4u,
                                              unsigned(c) * cs4 + unsigned(pbase) * 4u, 0);
    }
};

// h = m * [A Q - B (Q^2 - tr(Q^2)/3 I) + C tr(Q^2) Q]   (QSR.cu:399-415); the off-diagonal
// entries of Q^2 are reduced with tr Q = 0 (e.g. (Q^2)_xy = Q_xz Q_yz - Q_xy Q_zz)
__device__ __forceinline__ void bulk_field(const float q[5], float cA, float cB, float cC, float h[5]) {
    const float xx = q[0], xy = q[1], xz = q[2], yy = q[3], yz = q[4], zz = -(xx + yy);
    const float q2xx = __builtin_fmaf(xx, xx, __builtin_fmaf(xy, xy, xz * xz));
    const float q2yy = __builtin_fmaf(xy, xy, __builtin_fmaf(yy, yy, yz * yz));
    const float q2zz = __builtin_fmaf(xz, xz, __builtin_fmaf(yz, yz, zz * zz));
    const float tr2 = q2xx + q2yy + q2zz;
    const float t3 = tr2 * (1.0f / 3.0f);
    const float g = __builtin_fmaf(cC, tr2, cA);
    h[0] = __builtin_fmaf(g, xx, -cB * (q2xx - t3));
    h[1] = __builtin_fmaf(g, xy, -cB * __builtin_fmaf(xz, yz, -(xy * zz)));
    h[2] = __builtin_fmaf(g, xz, -cB * __builtin_fmaf(xy, yz, -(xz * yy)));
    h[3] = __builtin_fmaf(g, yy, -cB * (q2yy - t3));
    h[4] = __builtin_fmaf(g, yz, -cB * __builtin_fmaf(xy, xz, -(yz * xx)));
}

// Q0 = S (nu nu - I/3), nu = normalised (x/rx^2, y/ry^2, z/rz^2)  (QSR.cu:60-68, 130-142)
__device__ __forceinline__ void shell_target(const BoxDev& B, int i, int j, int k, float S, float q0[5]) {
    float nx = (float(i) - B.ci) * B.sx, ny = (float(j) - B.cj) * B.sy, nz = (float(k) - B.ck) * B.sz;
    const float n2 = nx * nx + ny * ny + nz * nz;
    if (n2 > 1e-30f) {
        const float inv = rsqrtf(n2);
        nx *= inv; ny *= inv; nz *= inv;
    } else {
        nx = 1.0f; ny = 0.0f; nz = 0.0f;
    }
    q0[0] = S * (nx * nx - (1.0f / 3.0f));
    q0[1] = S * (nx * ny);
    q0[2] = S * (nx * nz);
    q0[3] = S * (ny * ny - (1.0f / 3.0f));
    q0[4] = S * (ny * nz);
}

// Philox4x32-10 (Salmon et al., SC'11)
__device__ __forceinline__ void philox10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32->64 product per multiplier (a single v_mad_u64_u32 each)
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
        const unsigned n0 = unsigned(p1 >> 32) ^ c[1] ^ k0, n2 = unsigned(p0 >> 32) ^ c[3] ^ k1;
        c[0] = n0; c[1] = unsigned(p1); c[2] = n2; c[3] = unsigned(p0);
    }
}

// Five N(0,1) variates keyed by (seed, voxel, iteration): one Philox4x32-10
// block with counter (voxel lo, voxel hi, iteration, 0) and key (seed lo,
// seed hi); its 128 bits are cut into six 21-bit fields f_t = bits
// [21t, 21t + 21), turned into uniforms u_t = (f_t + 1/2) 2^-21 (exact in
// binary32), and paired by Box-Muller: sqrt(-2 ln u_2p) (cos, sin)(2 pi u_2p+1).
// The transform uses the hardware log2, sqrt and revolution-input sin/cos.
__device__ __forceinline__ unsigned field21(unsigned lo, unsigned hi, int shift) {
    return unsigned(((((unsigned long long)hi) << 32) | lo) >> shift) & 0x1FFFFFu;
}

__device__ __forceinline__ void noise5(unsigned long long seed, unsigned long long voxel, unsigned iter, float g[5]) {
    unsigned r[4] = {unsigned(voxel), unsigned(voxel >> 32), iter, 0u};
    philox10(r, unsigned(seed), unsigned(seed >> 32));
    const unsigned f[6] = {r[0] & 0x1FFFFFu,      field21(r[0], r[1], 21), (r[1] >> 10) & 0x1FFFFFu,
                           field21(r[1], r[2], 31), field21(r[2], r[3], 20), (r[3] >> 9) & 0x1FFFFFu};
    constexpr float kScale = 1.0f / 2097152.0f;       // 2^-21
    constexpr float kM2Ln2 = -1.3862943611198906f;    // -2 ln 2
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const float u0 = (float(f[2 * p]) + 0.5f) * kScale, u1 = (float(f[2 * p + 1]) + 0.5f) * kScale;
        const float rad = __builtin_amdgcn_sqrtf(kM2Ln2 * __builtin_amdgcn_logf(u0));
        g[2 * p] = rad * __builtin_amdgcn_cosf(u1);
        if (2 * p + 1 < 5) g[2 * p + 1] = rad * __builtin_amdgcn_sinf(u1);
    }
}

// Shared tail: anchoring penalty, Euler update, noise, cap, strong BC.
// h is m-scaled and already traceless.  Returns false if the voxel is static.
template <bool NOISE>
__device__ __forceinline__ bool finish_voxel(const BoxDev& B, const StepDyn& D, int i, int j, int k,
                                             long long local, bool evolve, bool shell, const float q[5],
                                             float h[5], float out[5], bool rhs_only = false) {
    const int fl = B.flags;
    const bool update = evolve && !((fl & F_SKIPSHELL) && shell);
    if (rhs_only) {
        // semi-implicit right-hand side (QSR.cu:734-784): frozen voxels are never read by the sweeps
        if (!update) return false;
        float q0a[5];
        if ((fl & F_ANCHOR) && shell) {
            shell_target(B, i, j, k, D.S_shell, q0a);
#pragma unroll
            for (int c = 0; c < 5; ++c) h[c] += B.cW * (q[c] - q0a[c]);
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = q[c] - h[c];
        return true;
    }
    const bool bc = (fl & F_BC) && shell;
    if (!update && !bc) return false;
    float q0[5];
    if (shell && (fl & (F_ANCHOR | F_BC))) shell_target(B, i, j, k, D.S_shell, q0);
    if ((fl & F_ANCHOR) && shell) {
#pragma unroll
        for (int c = 0; c < 5; ++c) h[c] += B.cW * (q[c] - q0[c]);
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = q[c];
    if (update) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = q[c] - h[c];
        if (NOISE && (fl & F_NOISE)) {
            float g[5];
            noise5(B.seed, (unsigned long long)local, D.iter, g);
#pragma unroll
            for (int c = 0; c < 5; ++c) out[c] += B.sigma * g[c];
        }
        if (fl & F_CAP) {
            const float xx = out[0], yy = out[3], zz = -(xx + yy);
            const float n2 = xx * xx + yy * yy + zz * zz + 2.0f * (out[1] * out[1] + out[2] * out[2] + out[4] * out[4]);
            if (n2 > B.cap2 && n2 > 0.0f) {
                const float s = sqrtf(B.cap2 / n2);
#pragma unroll
                for (int c = 0; c < 5; ++c) out[c] *= s;
            }
        }
    }
    if (bc) {
#pragma unroll
        for (int c = 0; c < 5; ++c) out[c] = q0[c];
    }
    return true;
}

__device__ __forceinline__ void decode_row(const RowMask& m, const BoxDev& B, int k, bool& evolve, bool& shell) {
    if (B.flags & F_PERIODIC) {
        evolve = true;
        shell = false;
        return;
    }
    evolve = in_iv(k, m.v[0], m.v[1]);
    shell = in_iv(k, m.v[2], m.v[3]) && !in_iv(k, m.v[4], m.v[5]);
}

__device__ __forceinline__ void decode_mask(const RowMask* rows, const BoxDev& B, int i, int j, int k,
                                            bool& evolve, bool& shell) {
    RowMask m = {};
    if (!(B.flags & F_PERIODIC)) m = rows[B.rows_off + (long long)i * B.Ny + j];
    decode_row(m, B, k, evolve, shell);
}

// Row-mask record of row (i, j) as raw 16 bytes, for prefetching a plane ahead:
// kept packed so the conditional load has no consumer before the loop moves on
// (unpacking right after the load would force a wait for it, and with it for
// every prefetch issued earlier).  Callers only pass rows inside a clamped box.
__device__ __forceinline__ int4 load_row4(const RowMask* rows, const BoxDev& B, int i, int j) {
    return *reinterpret_cast<const int4*>(rows + B.rows_off + (long long)i * B.Ny + j);
}

// The first 12 bytes of the record (the fourth word is padding), for kernels
// short of registers; decoded by the same decode_row4.
struct Row3 {
    int x, y, z;
};
__device__ __forceinline__ Row3 load_row3(const RowMask* rows, const BoxDev& B, int i, int j) {
    const int* r = reinterpret_cast<const int*>(rows + B.rows_off + (long long)i * B.Ny + j);
    return Row3{r[0], r[1], r[2]};
}

template <class M>
__device__ __forceinline__ void decode_row4(M m, const BoxDev& B, int k, bool& evolve, bool& shell) {
    if (B.flags & F_PERIODIC) {
        evolve = true;
        shell = false;
        return;
    }
    const int e0 = (short)(m.x & 0xffff), e1 = m.x >> 16, o0 = (short)(m.y & 0xffff), o1 = m.y >> 16;
    const int n0 = (short)(m.z & 0xffff), n1 = m.z >> 16;
    evolve = k >= e0 && k <= e1;
    shell = (k >= o0 && k <= o1) && !(k >= n0 && k <= n1);
}

}  // namespace

// L1 / kappa molecular field of one voxel, m-scaled: bulk minus L1 lap Q (QSR.cu:279-323,
// 399-415), from the neighbour sums xs = Q(i-1) + Q(i+1), ys, zs along y and z:
// L1 lap Q = lx xs + ly ys + lz zs - l2 Q with l2 = 2 (lx + ly + lz).  The single-step and
// two-step kernels both call this, so they round every voxel alike.
__device__ __forceinline__ void l1_field(const BoxDev& B, const StepDyn& D, float l2, const float qc[5],
                                         const float xs[5], const float ys[5], const float zs[5], float h[5]) {
    bulk_field(qc, D.cA, B.cB, B.cC, h);
#pragma unroll
    for (int c = 0; c < 5; ++c)
        h[c] = __builtin_fmaf(l2, qc[c],
                              __builtin_fmaf(-B.lz, zs[c], __builtin_fmaf(-B.ly, ys[c], __builtin_fmaf(-B.lx, xs[c], h[c]))));
}

// ---------------------------------------------------------------------------
// L1 / kappa kernel: x-register queue + one LDS plane (halo 1), double-buffered.
// ---------------------------------------------------------------------------
// Without noise the kernel fits 7 waves per SIMD (72 VGPRs, no spills; 6 waves uncapped): the
// step is latency bound, so the extra wave pays.  The noisy variant (Philox + Box-Muller) keeps
// its registers (5 waves); so do 64-bit-offset contexts, which would spill under the cap.
// reverse != 0 takes the tile list back to front; launches alternate it, so a launch first
// reads the tiles the previous one wrote last (tiles are independent: no value changes).
template <int TK, int TJ, bool NOISE, bool BUF>
__global__ __launch_bounds__(TK* TJ, (!NOISE && BUF) ? 7 : 1) void step_l1_kernel(const float* __restrict__ qin, float* __restrict__ qout,
                                                          long long cs, const BoxDev* __restrict__ boxes,
                                                          const int4* __restrict__ tiles,
                                                          const StepDyn* __restrict__ dyn, int nboxes,
                                                          const RowMask* __restrict__ rows, int ch,
                                                          const float4* __restrict__ imex_tab, int reverse) {
    constexpr int W = TK + 2, R = TJ + 2, PL = W * R, SLOT = 5 * PL;
    __shared__ float lds[2 * SLOT];
    const int4 t = tiles[reverse ? gridDim.x - 1 - blockIdx.x : blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD-ordered list (uniform over the block)
    const BoxDev B = boxes[t.x];
    const StepDyn D = dyn[t.x];
    const float4 imex = imex_tab ? imex_tab[t.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float l2 = uni(2.0f * (B.lx + B.ly + B.lz));
    const int tk = threadIdx.x & (TK - 1), tj = threadIdx.x / TK;
    const int Nx = B.Nx, Ny = B.Ny, Nz = B.Nz;
    const bool per = (B.flags & F_PERIODIC) != 0;
    int ib = t.y, ie = min(t.y + ch, Nx);
    if (!per) {
        ib = max(ib, 1);
        ie = min(ie, Nx - 1);
    }
    if (ib >= ie) return;  // uniform over the block
    const int j = t.z + tj, k = t.w + tk;
    const bool out_ok = per ? (j < Ny && k < Nz) : (j >= 1 && j <= Ny - 2 && k >= 1 && k <= Nz - 2);
    const int jc = wrap_or_clamp(j, Ny, per), kc = wrap_or_clamp(k, Nz, per);
    // Halo ring of the plane (rows j0-1 and j0+TJ, columns k0-1 and k0+TK, no
    // corners): thread tid < 2 TK + 2 TJ owns ring point tid, so a thread holds at
    // most one halo point and the row halos are loaded by whole coalesced waves.
    const int tid = threadIdx.x;
    const bool halo = tid < 2 * TK + 2 * TJ;
    int hj = 0, hk = 0;  // ring point in tile coordinates (-1 .. TJ, -1 .. TK)
    if (tid < TK) { hj = -1; hk = tid; }
    else if (tid < 2 * TK) { hj = TJ; hk = tid - TK; }
    else if (tid < 2 * TK + TJ) { hj = tid - 2 * TK; hk = -1; }
    else if (halo) { hj = tid - 2 * TK - TJ; hk = TK; }
    const long long plane = (long long)Ny * Nz;
    const long long o_own = (long long)jc * Nz + kc;
    const long long o_halo = (long long)wrap_or_clamp(t.z + hj, Ny, per) * Nz + wrap_or_clamp(t.w + hk, Nz, per);
    const Field<BUF> F(qin, qout, cs, B.off);

    auto ld5 = [&](int p, long long o, float v[5]) {
        const long long pb = (long long)p * plane;
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = F.ld(c, pb, o);
    };
    // own queue Q(p-1), Q(p), Q(p+1) plus Q(p+2) in flight: every global load is
    // consumed one whole iteration after it is issued
    float qm[5], qc[5], qn[5], hc[5] = {}, hn[5] = {};
    ld5(wrap_or_clamp(ib - 1, Nx, per), o_own, qm);
    ld5(ib, o_own, qc);
    ld5(wrap_or_clamp(ib + 1, Nx, per), o_own, qn);
    if (halo) ld5(ib, o_halo, hc);
    const int lrow = (tj + 1) * W + tk + 1;
    const int lhalo = (hj + 1) * W + hk + 1;

    const bool pre_rows = out_ok && !per;  // rows with a mask record to prefetch
    Row3 mcur{0, 0, 0};
    if (pre_rows) mcur = load_row3(rows, B, ib, j);
    for (int p = ib; p < ie; ++p) {
        // next plane's halo and row mask, and own Q(p+2), a whole iteration ahead of use
        const bool more = p + 1 < ie;
        const int pn = wrap_or_clamp(p + 1, Nx, per);
        Row3 mnext{0, 0, 0};
        float qf[5];
        if (more) {
            if (pre_rows) mnext = load_row3(rows, B, p + 1, j);
            ld5(wrap_or_clamp(p + 2, Nx, per), o_own, qf);
            if (halo) ld5(pn, o_halo, hn);
        }
        float* L = lds + (p & 1) * SLOT;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            L[c * PL + lrow] = qc[c];
            if (halo) L[c * PL + lhalo] = hc[c];
        }
        lds_barrier();
        if (out_ok) {
            bool evolve, shell;
            decode_row4(mcur, B, k, evolve, shell);
            const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
            if (live) {
                float h[5], xs[5], ys[5], zs[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const float* Lc = L + c * PL;
                    xs[c] = qn[c] + qm[c];
                    ys[c] = Lc[lrow + W] + Lc[lrow - W];
                    zs[c] = Lc[lrow + 1] + Lc[lrow - 1];
                }
                l1_field(B, D, l2, qc, xs, ys, zs, h);
                if (imex.w != 0.0f) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const float q2 = 2.0f * qc[c];
                        h[c] += imex.x * (xs[c] - q2) + imex.y * (ys[c] - q2) + imex.z * (zs[c] - q2);
                    }
                }
                float o[5];
                const long long inpl = (long long)j * Nz + k, pb = (long long)p * plane;
                if (finish_voxel<NOISE>(B, D, p, j, k, pb + inpl, evolve, shell, qc, h, o, imex.w != 0.0f)) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) F.st(c, pb, inpl, o[c]);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            qm[c] = qc[c];
            qc[c] = qn[c];
            qn[c] = more ? qf[c] : qn[c];
            hc[c] = hn[c];
        }
        mcur = mnext;
    }
}

// ---------------------------------------------------------------------------
// Two L1 / kappa steps in one pass (temporal blocking): the launch reads the field
// once and writes it once for two explicit steps.  A block of 256 threads covers a
// region of 32 k x 8 j points and streams along i; step 1 updates every region
// point, step 2 the inner 30 x 6 points, which are the tile's outputs.  For each
// plane s of the stream:
//   * plane s of the input (own column + a one-point ring around the region) goes
//     into an LDS slot; the own column also runs through a register queue along i;
//   * step 1 gives the intermediate field at plane s on the whole region (its own
//     register queue along i and LDS slot);
//   * step 2 gives output plane s-1 on the inner points.
// Both LDS planes are double-buffered, so one barrier per plane orders every
// hand-over.  Usable when no box face can change (the face pass would have to run
// between the two steps), without noise and without the semi-implicit right-hand
// side; 32-bit buffer offsets (5 x cs floats below 4 GiB).
// The stores are ordinary cache-allocating ones, so the memory-side cache still
// holds much of the field when the next launch reads it.  reverse != 0 takes the
// tile list back to front: launches alternate it, so the next launch first reads
// the tiles written last, the ones the cache has kept.  Tiles are independent, so
// the order changes no value.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(X2_TK* X2_TJ, 5) void step_l1x2_kernel(
    const float* __restrict__ qin, float* __restrict__ qout, long long cs, const BoxDev* __restrict__ boxes,
    const int4* __restrict__ tiles, const StepDyn* __restrict__ dyn1, const StepDyn* __restrict__ dyn2,
    const RowMask* __restrict__ rows, int ch, int reverse) {
    constexpr int TK = X2_TK, TJ = X2_TJ, NT = TK * TJ;
    constexpr int WA = TK + 2, PA = WA * (TJ + 2), SA = 5 * PA;  // input plane with its ring
    constexpr int PB = NT, SB = 5 * PB;                          // intermediate plane (region)
    __shared__ float la[2 * SA];
    __shared__ float lb[2 * SB];
    const int4 t = tiles[reverse ? gridDim.x - 1 - blockIdx.x : blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD-ordered list (uniform over the block)
    const BoxDev B = boxes[t.x];
    const StepDyn D1 = dyn1[t.x], D2 = dyn2[t.x];
    const float l2 = uni(2.0f * (B.lx + B.ly + B.lz));
    const int tid = threadIdx.x, tk = tid & (TK - 1), tj = tid / TK;
    const int Nx = B.Nx, Ny = B.Ny, Nz = B.Nz;
    const bool per = (B.flags & F_PERIODIC) != 0;
    int ib = t.y, ie = min(t.y + ch, Nx);
    if (!per) {
        ib = max(ib, 1);
        ie = min(ie, Nx - 1);
    }
    if (ib >= ie) return;  // uniform over the block
    // region point of this thread (tile outputs start at (t.z, t.w), the region one point before)
    const int j = t.z - 1 + tj, k = t.w - 1 + tk;
    const bool inner = tj >= 1 && tj <= TJ - 2 && tk >= 1 && tk <= TK - 2;
    // points a step may change: the interior of a clamped box, all of a periodic one
    const bool jk_in = per ? true : (j >= 1 && j <= Ny - 2 && k >= 1 && k <= Nz - 2);
    const bool out_ok = inner && (per ? (j < Ny && k < Nz) : jk_in);
    const int jc = wrap_or_clamp(j, Ny, per), kc = wrap_or_clamp(k, Nz, per);
    // ring point of the region (rows t.z-2 and t.z-1+TJ, columns t.w-2 and t.w-1+TK), one per thread
    const bool ring = tid < 2 * TK + 2 * TJ;
    int hj = 0, hk = 0;  // region coordinates, -1 .. TJ and -1 .. TK
    if (tid < TK) { hj = -1; hk = tid; }
    else if (tid < 2 * TK) { hj = TJ; hk = tid - TK; }
    else if (tid < 2 * TK + TJ) { hj = tid - 2 * TK; hk = -1; }
    else if (ring) { hj = tid - 2 * TK - TJ; hk = TK; }
    const long long plane = (long long)Ny * Nz;
    const long long o_own = (long long)jc * Nz + kc;
    const long long o_ring = (long long)wrap_or_clamp(t.z - 1 + hj, Ny, per) * Nz + wrap_or_clamp(t.w - 1 + hk, Nz, per);
    const Field<true> F(qin, qout, cs, B.off);
    const int pa = (tj + 1) * WA + tk + 1, pr = (hj + 1) * WA + hk + 1, pb = tj * TK + tk;
    const int jrow = min(max(j, 0), Ny - 1);
    auto pl = [&](int p) { return (long long)wrap_or_clamp(p, Nx, per) * plane; };
    auto ld5 = [&](int p, long long o, float v[5]) {
        const long long b = pl(p);
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = F.ld(c, b, o);
    };
    auto row = [&](int p) { return load_row3(rows, B, min(max(p, 0), Nx - 1), jrow); };
    // one explicit step at this thread's point of plane p: qm, qc, qn along i, in-plane
    // neighbours from slot S (row stride W, own position pos); out = qc where nothing changes
    auto step_at = [&](const StepDyn& D, int p, Row3 m, const float qm[5], const float qc[5], const float qn[5],
                       const float* S, int PS, int W, int pos, float out[5]) -> bool {
        bool evolve, shell;
        decode_row4(m, B, k, evolve, shell);
        const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
        if (!live) return false;
        float h[5], xs[5], ys[5], zs[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float* Sc = S + c * PS;
            xs[c] = qn[c] + qm[c];
            ys[c] = Sc[pos + W] + Sc[pos - W];
            zs[c] = Sc[pos + 1] + Sc[pos - 1];
        }
        l1_field(B, D, l2, qc, xs, ys, zs, h);
        return finish_voxel<false>(B, D, p, j, k, 0, evolve, shell, qc, h, out);
    };

    // queues: input Q(s-1), Q(s), Q(s+1) (+ Q(s+2) in flight), intermediate Q'(s-2), Q'(s-1)
    const int s0 = ib - 1;
    float qm[5], qc[5], qn[5], hc[5] = {}, pm[5] = {}, pc[5] = {};
    ld5(s0 - 1, o_own, qm);
    ld5(s0, o_own, qc);
    ld5(s0 + 1, o_own, qn);
    if (ring) ld5(s0, o_ring, hc);
    Row3 ms = row(s0), mu{0, 0, 0};  // row masks of planes s (step 1) and s-1 (step 2)
    for (int s = s0; s <= ie; ++s) {
        const bool more = s < ie;
        float qf[5], hn[5] = {};
        Row3 mn{0, 0, 0};
        if (more) {
            ld5(s + 2, o_own, qf);
            if (ring) ld5(s + 1, o_ring, hn);
            mn = row(s + 1);
        }
        float* A = la + (s & 1) * SA;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            A[c * PA + pa] = qc[c];
            if (ring) A[c * PA + pr] = hc[c];
        }
        lds_barrier();
        // step 1 at plane s
        float p1[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) p1[c] = qc[c];
        if (jk_in && (per || (s >= 1 && s <= Nx - 2))) {
            float o[5];
            if (step_at(D1, s, ms, qm, qc, qn, A, PA, WA, pa, o)) {
#pragma unroll
                for (int c = 0; c < 5; ++c) p1[c] = o[c];
            }
        }
        float* Bs = lb + (s & 1) * SB;
#pragma unroll
        for (int c = 0; c < 5; ++c) Bs[c * PB + pb] = p1[c];
        // step 2 at plane s-1 from Q'(s-2), Q'(s-1) (slot written last plane) and Q'(s)
        const int u = s - 1;
        if (u >= ib && out_ok) {
            float o[5];
            if (step_at(D2, u, mu, pm, pc, p1, lb + (u & 1) * SB, PB, TK, pb, o)) {
                const long long inpl = (long long)j * Nz + k, ub = (long long)u * plane;
#pragma unroll
                for (int c = 0; c < 5; ++c) F.st_cached(c, ub, inpl, o[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            pm[c] = pc[c];
            pc[c] = p1[c];
            qm[c] = qc[c];
            qc[c] = qn[c];
            qn[c] = more ? qf[c] : qn[c];
            hc[c] = hn[c];
        }
        mu = ms;
        ms = mn;
    }
}

// ---------------------------------------------------------------------------
// Anisotropic kernel (L2 and/or L3).  Streams along i; output plane t is
// produced after lead plane t+1 has been processed.
//   * Three Q planes t, t+1 and (written at the end of the iteration) t+2 sit
//     in LDS with an in-plane halo of 2 (a ring of three slots, so plane t is
//     still resident while plane t+2 is written).  Every in-plane derivative
//     of the own point (d_y Q, d_z Q, second differences, mixed d_y d_z Q) and
//     the own values Q(t), Q(t+1) are read from the slots, not carried in
//     registers.
//   * Registers carry only Q(t-1) and Q(t+2) (plus the prefetch of Q(t+3)),
//     Dcol(t-1) and, for ring threads, Q(t+2) of their ring point, so the
//     kernel fits four waves per SIMD.
//   * x-derivatives: central differences of the register values; the mixed
//     d_x d_y Q = (d_yQ(t+1) - d_yQ(t-1)) / (2 dx) is the reference's 4-point
//     formula (QSR.cu:583-588).  The L3 term needs it only inside
//     Q_xy d_x d_y Q + Q_xz d_x d_z Q: its d_yQ/d_zQ(t-1) half is folded one
//     plane early (pxm), while plane t-1 is still resident, and its t+1 half
//     (pxp) while plane t+1's derivatives are formed.
//   * Dcol = d_k Q_kj (QSR.cu:325-371) in an LDS plane with a halo of 1
//     (two slots): own points reuse their first derivatives, the halo ring is
//     computed by ring threads.
// One barrier per plane orders every LDS hand-over (see the loop).
// ---------------------------------------------------------------------------
namespace {

template <int TK, int TJv>
struct AnisoGeom {
    static constexpr int W2 = TK + 4, R2 = TJv + 4, P2 = W2 * R2, QSLOT = 5 * P2;
    static constexpr int W1 = TK + 2, R1 = TJv + 2, P1 = W1 * R1, DSLOT = 3 * P1;
    static constexpr int NT = TK * TJv;
    static constexpr int NRING = 2 * TK + 2 * TJv;
    static constexpr int NHALO = P2 - TK * TJv;               // halo points of a lead plane
    static constexpr int NHPT = (NHALO + NT - 1) / NT;         // halo points per thread
};

}  // namespace

// EL: elastic terms compiled in (1 = L2, 2 = L3, 3 = both), the union over the
// context's boxes; a box without a term has a zero coefficient for it, which
// adds exactly nothing, so the whole molecular field is one basic block.
// Four waves per SIMD (128 VGPRs) for the instances the headline runs use; the
// noisy and 64-bit-offset instances keep three.
template <int TK, int TJv, bool NOISE, bool BUF, int EL>
__global__ __launch_bounds__(TK* TJv, (!NOISE && BUF) ? 4 : 3) void step_aniso_kernel(
    const float* __restrict__ qin, float* __restrict__ qout, long long cs, const BoxDev* __restrict__ boxes,
    const int4* __restrict__ tiles, const StepDyn* __restrict__ dyn, int nboxes, const RowMask* __restrict__ rows,
    int ch, const float4* __restrict__ imex_tab) {
    using G = AnisoGeom<TK, TJv>;
    constexpr int W2 = G::W2, P2 = G::P2, QSLOT = G::QSLOT, W1 = G::W1, P1 = G::P1, DSLOT = G::DSLOT;
    __shared__ float qs[3 * QSLOT];
    __shared__ float ds[2 * DSLOT];
    const int4 t4 = tiles[blockIdx.x];
    if (t4.x < 0) return;  // padding tile of the XCD-ordered list (uniform over the block)
    const BoxDev B = boxes[t4.x];
    const StepDyn D = dyn[t4.x];
    const float4 imex = imex_tab ? imex_tab[t4.x] : make_float4(0.f, 0.f, 0.f, 0.f);
    // The molecular field works on raw differences (Q[+1] - Q[-1], Q[+1] + Q[-1] - 2Q, the
    // 4-point mixed difference); the grid scales and m L2, m L3 sit in these per-box products.
    //   second differences: L1 (minus the IMEX stabiliser) + m L3 Q_aa / d_a^2
    const float lxp = uni(B.lx - imex.x), lyp = uni(B.ly - imex.y), lzp = uni(B.lz - imex.z);
    const float k3x = uni(B.mL3 * B.ix2), k3y = uni(B.mL3 * B.iy2), k3z = uni(B.mL3 * B.iz2);
    const float k3yz = uni(2.0f * B.mL3 * B.i4yz);                      // 2 Q_yz d_y d_z
    const float k3hx = uni(B.mL3 * B.hx), k3hy = uni(B.mL3 * B.hy), k3hz = uni(B.mL3 * B.hz);  // Dcol . grad
    const float kpy = uni(2.0f * B.hx * B.hy * B.mL3), kpz = uni(2.0f * B.hx * B.hz * B.mL3);  // 2 Q_xa d_x d_a
    const float a2x = uni(B.mL2 * B.hx), a2y = uni(B.mL2 * B.hy), a2z = uni(B.mL2 * B.hz);     // d Dcol
    const float kxx = uni(B.mL3 * B.hx * B.hx), kyy = uni(B.mL3 * B.hy * B.hy), kzz = uni(B.mL3 * B.hz * B.hz);
    const float kxy = uni(B.mL3 * B.hx * B.hy), kxz = uni(B.mL3 * B.hx * B.hz), kyz = uni(B.mL3 * B.hy * B.hz);
    const int tid = threadIdx.x, tk = tid & (TK - 1), tj = tid / TK;
    const int Nx = B.Nx, Ny = B.Ny, Nz = B.Nz;
    const bool per = (B.flags & F_PERIODIC) != 0;
    int ib = t4.y, ie = min(t4.y + ch, Nx);
    if (!per) {
        ib = max(ib, 1);
        ie = min(ie, Nx - 1);
    }
    if (ib >= ie) return;
    const int j0 = t4.z, k0 = t4.w;
    const int j = j0 + tj, k = k0 + tk;
    const bool out_ok = per ? (j < Ny && k < Nz) : (j >= 1 && j <= Ny - 2 && k >= 1 && k <= Nz - 2);
    const long long plane = (long long)Ny * Nz;
    const Field<BUF> F(qin, qout, cs, B.off);
    constexpr bool hasL2 = (EL & 1) != 0, hasL3 = (EL & 2) != 0;
    auto face = [&](int x, int n) { return !per && (x == 0 || x == n - 1); };

    // own column
    const long long o_own = (long long)wrap_or_clamp(j, Ny, per) * Nz + wrap_or_clamp(k, Nz, per);
    const float ay_own = face(j, Ny) ? B.fy : B.hy, az_own = face(k, Nz) ? B.fz : B.hz;
    // ring point of this thread (if any), in region coordinates (rj, rk)
    const bool ring = tid < G::NRING;
    int rj = 0, rk = 0;
    if (ring) {
        if (tid < TK) { rj = j0 - 1; rk = k0 + tid; }
        else if (tid < 2 * TK) { rj = j0 + TJv; rk = k0 + tid - TK; }
        else if (tid < 2 * TK + TJv) { rj = j0 + tid - 2 * TK; rk = k0 - 1; }
        else { rj = j0 + tid - 2 * TK - TJv; rk = k0 + TK; }
    }
    const long long o_ring = (long long)wrap_or_clamp(rj, Ny, per) * Nz + wrap_or_clamp(rk, Nz, per);
    const float ay_ring = face(rj, Ny) ? B.fy : B.hy, az_ring = face(rk, Nz) ? B.fz : B.hz;
    const int lq_ring = (rj - j0 + 2) * W2 + (rk - k0 + 2);   // position in a Q slot
    const int ld_ring = (rj - j0 + 1) * W1 + (rk - k0 + 1);   // position in a D slot
    const int lq = (tj + 2) * W2 + tk + 2, ld = (tj + 1) * W1 + tk + 1;
    // halo points of a lead plane handled by this thread (halo = Q slot minus the interior)
    long long hoff[G::NHPT];
    int hpos[G::NHPT];
#pragma unroll
    for (int s = 0; s < G::NHPT; ++s) {
        const int h = tid + s * G::NT;
        hpos[s] = -1;
        hoff[s] = 0;
        if (h < G::NHALO) {
            // enumerate the slot row-major and skip the interior block
            int id = h, r = 0, c = 0;
            const int top = 2 * W2;                       // rows 0, 1
            const int mid = TJv * 4;                      // rows 2..TJv+1: 4 halo points each
            if (id < top) { r = id / W2; c = id % W2; }
            else if (id < top + mid) { id -= top; r = 2 + id / 4; c = id % 4; c = c < 2 ? c : TK + c; }
            else { id -= top + mid; r = TJv + 2 + id / W2; c = id % W2; }
            hpos[s] = r * W2 + c;
            hoff[s] = (long long)wrap_or_clamp(j0 - 2 + r, Ny, per) * Nz + wrap_or_clamp(k0 - 2 + c, Nz, per);
        }
    }

    auto pl = [&](int p) { return (long long)wrap_or_clamp(p, Nx, per) * plane; };
    auto ld5 = [&](int p, long long o, float v[5]) {
        const long long pb = pl(p);
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = F.ld(c, pb, o);
    };
    auto ld3 = [&](int p, long long o, float v[3]) {
        const long long pb = pl(p);
#pragma unroll
        for (int c = 0; c < 3; ++c) v[c] = F.ld(c, pb, o);
    };
    // halo of plane p (this thread's share) into registers
    auto load_halo = [&](int p, float hv[G::NHPT][5]) {
        const long long pb = pl(p);
#pragma unroll
        for (int s = 0; s < G::NHPT; ++s)
            if (hpos[s] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) hv[s][c] = F.ld(c, pb, hoff[s]);
            }
    };
    // plane p lives in Q slot (p - ib + 1) mod 3
    auto qslot = [&](int p) { return qs + ((p - ib + 1) % 3) * QSLOT; };
    // write a plane (own value `own` + halo `hv`) into its Q slot
    auto write_slot = [&](int p, const float own[5], const float hv[G::NHPT][5]) {
        float* Q = qslot(p);
#pragma unroll
        for (int c = 0; c < 5; ++c) Q[c * P2 + lq] = own[c];
#pragma unroll
        for (int s = 0; s < G::NHPT; ++s)
            if (hpos[s] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) Q[c * P2 + hpos[s]] = hv[s][c];
            }
    };
    auto fill_slot = [&](int p, const float own[5]) {
        float hv[G::NHPT][5];
        load_halo(p, hv);
        write_slot(p, own, hv);
    };
    // own values of a resident plane
    auto own_q = [&](const float* Q, float v[5]) {
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = Q[c * P2 + lq];
    };
    // own raw first differences Q[+1] - Q[-1] of a resident plane along y and z
    auto own_r1 = [&](const float* Q, float ry[5], float rz[5]) {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const float* A = Q + c * P2;
            ry[c] = A[lq + W2] - A[lq - W2];
            rz[c] = A[lq + 1] - A[lq - 1];
        }
    };
    // Dcol from x-part (three comps xx, xy, xz differenced along x) and raw in-plane differences
    // scaled by the point's face-aware y and z factors.  The quantities formed both in the
    // prologue and in the loop (Dcol, the folded mixed terms) spell out their fused
    // multiply-adds, and an own point and the same point as another tile's ring point use one
    // sequence, so every call site rounds them alike and the result does not depend on where a
    // tile's planes start (planes per tile, tile shape).
    auto dcol_own = [&](const float xm[5], const float xp[5], float ax, const float ry[5], const float rz[5],
                        float Dv[3]) {
        Dv[0] = __builtin_fmaf(ay_own, ry[1], __builtin_fmaf(az_own, rz[2], ax * (xp[0] - xm[0])));
        Dv[1] = __builtin_fmaf(ay_own, ry[3], __builtin_fmaf(az_own, rz[4], ax * (xp[1] - xm[1])));
        Dv[2] = __builtin_fmaf(ay_own, ry[4], __builtin_fmaf(-az_own, rz[0] + rz[3], ax * (xp[2] - xm[2])));
    };
    // ring Dcol of resident plane Q; the x-difference uses the ring point's values of the plane
    // before (read from slot Qm) and after (xp)
    auto dcol_ring = [&](const float* Q, const float* Qm, const float xp[3], float ax, float Dv[3]) {
        auto ryv = [&](int c) { return Q[c * P2 + lq_ring + W2] - Q[c * P2 + lq_ring - W2]; };
        auto rzv = [&](int c) { return Q[c * P2 + lq_ring + 1] - Q[c * P2 + lq_ring - 1]; };
        Dv[0] = __builtin_fmaf(ay_ring, ryv(1), __builtin_fmaf(az_ring, rzv(2), ax * (xp[0] - Qm[0 * P2 + lq_ring])));
        Dv[1] = __builtin_fmaf(ay_ring, ryv(3), __builtin_fmaf(az_ring, rzv(4), ax * (xp[1] - Qm[1 * P2 + lq_ring])));
        Dv[2] = __builtin_fmaf(ay_ring, ryv(4),
                               __builtin_fmaf(-az_ring, rzv(0) + rzv(3), ax * (xp[2] - Qm[2 * P2 + lq_ring])));
    };
    // L3: the t-1 or t+1 half of plane t's 2 Q_xy d_x d_y Q + 2 Q_xz d_x d_z Q (m L3 and the grid
    // scales included), from plane t's Q_xy, Q_xz and the raw differences of the plane beside it
    auto fold_x = [&](float qxy, float qxz, const float ry[5], const float rz[5], float p[5]) {
        const float cy = qxy * kpy, cz = qxz * kpz;
#pragma unroll
        for (int c = 0; c < 5; ++c) p[c] = __builtin_fmaf(cz, rz[c], cy * ry[c]);
    };
    auto axs = [&](int p) { return face(p, Nx) ? B.fx : B.hx; };
    auto store_d = [&](int dslot, int pos, const float Dv[3]) {
        float* Dd = ds + dslot * DSLOT;
        Dd[0 * P1 + pos] = Dv[0];
        Dd[1 * P1 + pos] = Dv[1];
        Dd[2 * P1 + pos] = Dv[2];
    };

    // ---- prologue: planes ib-1, ib, ib+1 resident; Dcol(ib-1) own, Dcol(ib) own and ring --
    float q0[5], q3[5];          // own Q(t-1), Q(t+2)
    // L3 only: the t-1 half of plane t's mixed x-y and x-z terms (fold_x), folded one plane
    // early while plane t-1 is resident
    float pxm[5] = {};
    float D0[3];                 // own Dcol(t-1)
    float r2[3] = {};            // ring comps xx, xy, xz at t+2
    {
        float qm2[5], q1[5], q2[5];
        ld5(ib - 2, o_own, qm2);
        ld5(ib - 1, o_own, q0);
        ld5(ib, o_own, q1);
        ld5(ib + 1, o_own, q2);
        ld5(ib + 2, o_own, q3);
        float rp[3] = {};
        if (ring) {
            ld3(ib + 1, o_ring, rp);
            ld3(ib + 2, o_ring, r2);
        }
        fill_slot(ib - 1, q0);
        fill_slot(ib, q1);
        fill_slot(ib + 1, q2);
        lds_barrier();
        {
            float rym[5], rzm[5];
            own_r1(qslot(ib - 1), rym, rzm);
            dcol_own(qm2, q1, axs(ib - 1), rym, rzm, D0);
            if constexpr (hasL3) fold_x(q1[1], q1[2], rym, rzm, pxm);
        }
        float ry1[5], rz1[5], D1[3];
        own_r1(qslot(ib), ry1, rz1);
        dcol_own(q0, q2, axs(ib), ry1, rz1, D1);
        store_d(ib & 1, ld, D1);
        if (ring) {
            float Dr[3];
            dcol_ring(qslot(ib), qslot(ib - 1), rp, axs(ib), Dr);
            store_d(ib & 1, ld_ring, Dr);
        }
        lds_barrier();
    }

    const bool pre_rows = out_ok && !per;  // rows with a mask record to prefetch
    int4 mcur = make_int4(0, 0, 0, 0);
    if (pre_rows) mcur = load_row4(rows, B, ib, j);
    float hv[G::NHPT][5];  // halo of the next lead plane to write, loaded an iteration ahead
    if (ib + 1 < ie) load_halo(ib + 2, hv);
    for (int t = ib; t < ie; ++t) {
        // LDS hand-over, one barrier per plane (at the end of the iteration):
        //   Q slot of plane t+2 (written in (e)) last held plane t-1, last read before the
        //   previous barrier; planes t and t+1 (read here) were written before it.
        //   D slot (t+1)&1 (written in (b)) was last read in (d) of iteration t-1; D slot t&1
        //   (read in (d)) was written in (b) of iteration t-1.
        const float* Qt = qslot(t);
        const float* Qn = qslot(t + 1);
        const int dsl = (t + 1) & 1;   // D slot of Dcol(t+1) (Dcol(L) sits in slot L & 1)
        const bool more = t + 1 < ie;
        // (b) lead plane t+1: raw first differences, own and ring Dcol(t+1), and (L3) the t+1 half
        // of plane t's mixed x-y and x-z terms
        float D2[3], pxp[5] = {};
        {
            float q1[5], ry2[5], rz2[5];
            own_q(Qt, q1);
            own_r1(Qn, ry2, rz2);
            dcol_own(q1, q3, axs(t + 1), ry2, rz2, D2);
            if constexpr (hasL3) fold_x(q1[1], q1[2], ry2, rz2, pxp);
        }
        store_d(dsl, ld, D2);
        if (ring) {
            float Dr[3];
            dcol_ring(Qn, Qt, r2, axs(t + 1), Dr);
            store_d(dsl, ld_ring, Dr);
        }
        // (e) lead plane t+2 into the slot that held plane t-1, which nothing reads after the
        // previous barrier (its own value is q3, its halo arrived with the previous iteration's
        // loads), so the halo registers are free before the output pass
        if (more) write_slot(t + 2, q3, hv);
        // the scheduling fence keeps the compiler from interleaving (b) and (d), which would
        // hold both passes' values at once
        __builtin_amdgcn_sched_barrier(0);
        // plane t's raw first differences: the output pass and the fold below both need them
        float ry1[5], rz1[5];
        own_r1(Qt, ry1, rz1);
        // (d) output plane t
        if (out_ok) {
            bool evolve, shell;
            decode_row4(mcur, B, k, evolve, shell);
            const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
            if (live) {
                const float* Dc = ds + (t & 1) * DSLOT;
                float q1[5], q2[5], sxx[5], syy[5], szz[5];
                own_q(Qt, q1);
                own_q(Qn, q2);
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const float* A = Qt + c * P2;
                    sxx[c] = __builtin_fmaf(-2.0f, q1[c], q2[c] + q0[c]);
                    syy[c] = __builtin_fmaf(-2.0f, q1[c], A[lq + W2] + A[lq - W2]);
                    szz[c] = __builtin_fmaf(-2.0f, q1[c], A[lq + 1] + A[lq - 1]);
                }
                // bulk, m [A Q - B (Q^2 - tr(Q^2)/3 I) + C tr(Q^2) Q] (QSR.cu:399-415), with the
                // off-diagonal entries of Q^2 reduced by tr Q = 0
                const float xx = q1[0], xy = q1[1], xz = q1[2], yy = q1[3], yz = q1[4], zz = -(xx + yy);
                const float q2xx = __builtin_fmaf(xx, xx, __builtin_fmaf(xy, xy, xz * xz));
                const float q2yy = __builtin_fmaf(xy, xy, __builtin_fmaf(yy, yy, yz * yz));
                const float q2zz = __builtin_fmaf(xz, xz, __builtin_fmaf(yz, yz, zz * zz));
                const float g = __builtin_fmaf(B.cC, q2xx + q2yy + q2zz, D.cA);
                // diagonal entries before the one trace projection of bulk, L2 and L3 (px, py, pz),
                // off-diagonal entries of h (h1, h2, h4)
                float px = -B.cB * q2xx, py = -B.cB * q2yy, pz = -B.cB * q2zz;
                float h1 = __builtin_fmaf(g, xy, -B.cB * __builtin_fmaf(xz, yz, -(xy * zz)));
                float h2 = __builtin_fmaf(g, xz, -B.cB * __builtin_fmaf(xy, yz, -(xz * yy)));
                float h4 = __builtin_fmaf(g, yz, -B.cB * __builtin_fmaf(xy, xz, -(yz * xx)));
                // per component: L1 lap Q (+ IMEX) and L3's Q_ab d_a d_b Q + Dcol . grad Q
                float e[5];
                if constexpr (hasL3) {
                    const float ex = __builtin_fmaf(k3x, xx, lxp), ey = __builtin_fmaf(k3y, yy, lyp);
                    const float ez = __builtin_fmaf(k3z, zz, lzp), eyz = k3yz * yz;
                    const float Dx = k3hx * Dc[0 * P1 + ld], Dy = k3hy * Dc[1 * P1 + ld], Dz = k3hz * Dc[2 * P1 + ld];
                    float rx[5];
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const float* A = Qt + c * P2;
                        rx[c] = q2[c] - q0[c];
                        const float syz = (A[lq + W2 + 1] - A[lq + W2 - 1]) - (A[lq - W2 + 1] - A[lq - W2 - 1]);
                        float v = __builtin_fmaf(Dz, rz1[c], pxp[c] - pxm[c]);
                        v = __builtin_fmaf(Dy, ry1[c], v);
                        v = __builtin_fmaf(Dx, rx[c], v);
                        v = __builtin_fmaf(eyz, syz, v);
                        v = __builtin_fmaf(ez, szz[c], v);
                        v = __builtin_fmaf(ey, syy[c], v);
                        e[c] = __builtin_fmaf(ex, sxx[c], v);
                    }
                    // (1/2) m L3 d_a Q_kl d_b Q_kl = m L3 h_a h_b N_ab over raw differences r, with
                    // N_ab = s_a s_b + w_a w_b + sum over xy, xz, yz of r_a r_b, s = r_xx + r_yy / 2,
                    // w = (sqrt(3) / 2) r_yy: the xx, yy, zz = -(xx + yy) terms are 2 s s + (3/2) r_yy r_yy
                    auto r = [&](int c, int a) { return a == 0 ? rx[c] : (a == 1 ? ry1[c] : rz1[c]); };
                    float sv[3], wv[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        sv[a] = __builtin_fmaf(0.5f, r(3, a), r(0, a));
                        wv[a] = 0.8660254037844386f * r(3, a);
                    }
                    auto N = [&](int a, int b) {
                        float v = r(4, a) * r(4, b);
                        v = __builtin_fmaf(r(2, a), r(2, b), v);
                        v = __builtin_fmaf(r(1, a), r(1, b), v);
                        v = __builtin_fmaf(wv[a], wv[b], v);
                        return __builtin_fmaf(sv[a], sv[b], v);
                    };
                    px = __builtin_fmaf(kxx, N(0, 0), px);
                    py = __builtin_fmaf(kyy, N(1, 1), py);
                    pz = __builtin_fmaf(kzz, N(2, 2), pz);
                    h1 = __builtin_fmaf(kxy, N(0, 1), h1);
                    h2 = __builtin_fmaf(kxz, N(0, 2), h2);
                    h4 = __builtin_fmaf(kyz, N(1, 2), h4);
                } else {
#pragma unroll
                    for (int c = 0; c < 5; ++c)
                        e[c] = __builtin_fmaf(lxp, sxx[c], __builtin_fmaf(lyp, syy[c], lzp * szz[c]));
                }
                if constexpr (hasL2) {
                    // m L2 sym d(Dcol): central differences of Dcol(t-1, t+1) and of plane t's D slot
                    const float* Dy0 = Dc + ld;
                    px = __builtin_fmaf(-a2x, D2[0] - D0[0], px);
                    py = __builtin_fmaf(-a2y, Dy0[1 * P1 + W1] - Dy0[1 * P1 - W1], py);
                    pz = __builtin_fmaf(-a2z, Dy0[2 * P1 + 1] - Dy0[2 * P1 - 1], pz);
                    h1 = __builtin_fmaf(-a2x, D2[1] - D0[1], __builtin_fmaf(-a2y, Dy0[W1] - Dy0[-W1], h1));
                    h2 = __builtin_fmaf(-a2x, D2[2] - D0[2], __builtin_fmaf(-a2z, Dy0[1] - Dy0[-1], h2));
                    h4 = __builtin_fmaf(-a2y, Dy0[2 * P1 + W1] - Dy0[2 * P1 - W1],
                                        __builtin_fmaf(-a2z, Dy0[1 * P1 + 1] - Dy0[1 * P1 - 1], h4));
                }
                const float tr = (px + py + pz) * (1.0f / 3.0f);
                float h[5];
                h[0] = __builtin_fmaf(g, xx, px - tr) - e[0];
                h[3] = __builtin_fmaf(g, yy, py - tr) - e[3];
                h[1] = h1 - e[1];
                h[2] = h2 - e[2];
                h[4] = h4 - e[4];
                float o[5];
                const long long inpl = (long long)j * Nz + k, pb = (long long)t * plane;
                if (finish_voxel<NOISE>(B, D, t, j, k, pb + inpl, evolve, shell, q1, h, o, imex.w != 0.0f)) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) F.st(c, pb, inpl, o[c]);
                }
            }
        }
        // from here to the barrier the wave runs at raised priority, so a wave done with its output
        // pass issues the next plane's loads and reaches the barrier ahead of others' arithmetic
        // (about 2 % per launch at 256^3; the results are the same)
        __builtin_amdgcn_s_setprio(2);
        // the t-1 half of plane t+1's mixed terms: Q(t+1) and the differences of plane t, both resident
        if constexpr (hasL3) fold_x(Qn[1 * P2 + lq], Qn[2 * P2 + lq], ry1, rz1, pxm);
        // loads for the next iteration: row mask of plane t+1, own and ring Q(t+3), halo of
        // lead plane t+3
        int4 mnext = make_int4(0, 0, 0, 0);
        float q4[5], r3[3] = {};
        const bool more2 = t + 2 < ie;
        if (more) {
            if (pre_rows) mnext = load_row4(rows, B, t + 1, j);
            ld5(t + 3, o_own, q4);
            if (ring) ld3(t + 3, o_ring, r3);
            if (more2) load_halo(t + 3, hv);
        }
        // advance: t+1 becomes t
        own_q(Qt, q0);
#pragma unroll
        for (int a = 0; a < 3; ++a) D0[a] = ds[(t & 1) * DSLOT + a * P1 + ld];
#pragma unroll
        for (int c = 0; c < 5; ++c) q3[c] = more ? q4[c] : q3[c];
#pragma unroll
        for (int a = 0; a < 3; ++a) r2[a] = r3[a];
        mcur = mnext;
        lds_barrier();
        __builtin_amdgcn_s_setprio(0);
    }
}

// ---------------------------------------------------------------------------
// Face kernel: box-face voxels of clamped boxes with the reference's one-sided
// differences (QSR.cu:245-273, 296-320, 444-452); L3 is interior-only
// (QSR.cu:484) so it does not appear here.  Direct global loads: the faces are
// a few percent of the box and mostly static (outside the droplet).
// ---------------------------------------------------------------------------
namespace {

struct FaceCtx {
    const float* base;
    long long cs, plane;
    int Nx, Ny, Nz;
    __device__ float q(int c, int i, int j, int k) const { return base[c * cs + i * plane + (long long)j * Nz + k]; }
    // get_Q_deriv semantics along one axis
    __device__ float dq(int c, int i, int j, int k, int axis, float half, float full) const {
        int n = axis == 0 ? Nx : (axis == 1 ? Ny : Nz);
        int x = axis == 0 ? i : (axis == 1 ? j : k);
        int ip = x, im = x;
        float s;
        if (x > 0 && x < n - 1) { ip = x + 1; im = x - 1; s = half; }
        else if (x == 0) { ip = 1; im = 0; s = full; }
        else { ip = n - 1; im = n - 2; s = full; }
        auto at = [&](int v) {
            return axis == 0 ? q(c, v, j, k) : (axis == 1 ? q(c, i, v, k) : q(c, i, j, v));
        };
        return (at(ip) - at(im)) * s;
    }
    __device__ void dcol(const BoxDev& B, int i, int j, int k, float D[3]) const {
        const float xx_x = dq(0, i, j, k, 0, B.hx, B.fx), xy_y = dq(1, i, j, k, 1, B.hy, B.fy), xz_z = dq(2, i, j, k, 2, B.hz, B.fz);
        const float xy_x = dq(1, i, j, k, 0, B.hx, B.fx), yy_y = dq(3, i, j, k, 1, B.hy, B.fy), yz_z = dq(4, i, j, k, 2, B.hz, B.fz);
        const float xz_x = dq(2, i, j, k, 0, B.hx, B.fx), yz_y = dq(4, i, j, k, 1, B.hy, B.fy);
        const float zz_z = -(dq(0, i, j, k, 2, B.hz, B.fz) + dq(3, i, j, k, 2, B.hz, B.fz));
        D[0] = xx_x + xy_y + xz_z;
        D[1] = xy_x + yy_y + yz_z;
        D[2] = xz_x + yz_y + zz_z;
    }
};


}  // namespace

__global__ __launch_bounds__(256) void step_face_kernel(const float* __restrict__ qin, float* __restrict__ qout,
                                                        long long cs, const BoxDev* __restrict__ boxes,
                                                        const StepDyn* __restrict__ dyn,
                                                        const int4* __restrict__ face_list, long long nlive,
                                                        const RowMask* __restrict__ rows) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nlive) return;
    const int4 e = face_list[g];  // (box, i, j, k) of a face voxel that can change
    const int b = e.x, i = e.y, j = e.z, k = e.w;
    const BoxDev B = boxes[b];
    if (B.flags & (F_PERIODIC | F_INACTIVE)) return;
    const StepDyn D = dyn[b];
    bool evolve, shell;
    decode_mask(rows, B, i, j, k, evolve, shell);
    const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
    if (!live) return;
    FaceCtx F{qin + B.off, cs, (long long)B.Ny * B.Nz, B.Nx, B.Ny, B.Nz};
    float q[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) q[c] = F.q(c, i, j, k);
    float h[5];
    bulk_field(q, D.cA, B.cB, B.cC, h);
    // one-sided second difference on faces = interior stencil at the clamped index
    const int mi = min(max(i, 1), B.Nx - 2), mj = min(max(j, 1), B.Ny - 2), mk = min(max(k, 1), B.Nz - 2);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const float s_x = F.q(c, mi + 1, j, k) - 2.0f * F.q(c, mi, j, k) + F.q(c, mi - 1, j, k);
        const float s_y = F.q(c, i, mj + 1, k) - 2.0f * F.q(c, i, mj, k) + F.q(c, i, mj - 1, k);
        const float s_z = F.q(c, i, j, mk + 1) - 2.0f * F.q(c, i, j, mk) + F.q(c, i, j, mk - 1);
        h[c] -= B.lx * s_x + B.ly * s_y + B.lz * s_z;
    }
    if (B.flags & F_L2) {
        // d(Dcol) with get_deriv semantics (QSR.cu:245-253) at this face voxel
        float Dc[3];
        F.dcol(B, i, j, k, Dc);
        float dD[3][3];  // dD[axis][component]
        const int coord[3] = {i, j, k};
        const int nn[3] = {B.Nx, B.Ny, B.Nz};
        const float half[3] = {B.hx, B.hy, B.hz}, full[3] = {B.fx, B.fy, B.fz};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int x = coord[a], n = nn[a];
            float Dp[3], Dm[3], s;
            int pp[3] = {i, j, k}, mm[3] = {i, j, k};
            if (x > 0 && x < n - 1) { pp[a] = x + 1; mm[a] = x - 1; s = half[a]; }
            else if (x == 0) { pp[a] = 1; mm[a] = 0; s = full[a]; }
            else { pp[a] = n - 1; mm[a] = n - 2; s = full[a]; }
            F.dcol(B, pp[0], pp[1], pp[2], Dp);
            F.dcol(B, mm[0], mm[1], mm[2], Dm);
#pragma unroll
            for (int c = 0; c < 3; ++c) dD[a][c] = (Dp[c] - Dm[c]) * s;
        }
        const float gxx = B.mL2 * dD[0][0], gyy = B.mL2 * dD[1][1], gzz = B.mL2 * dD[2][2];
        const float tr = (gxx + gyy + gzz) * (1.0f / 3.0f);
        h[0] -= gxx - tr;
        h[3] -= gyy - tr;
        h[1] -= B.mL2 * (dD[0][1] + dD[1][0]);
        h[2] -= B.mL2 * (dD[0][2] + dD[2][0]);
        h[4] -= B.mL2 * (dD[1][2] + dD[2][1]);
    }
    float o[5];
    const long long local = (long long)i * F.plane + (long long)j * B.Nz + k;
    if (finish_voxel<true>(B, D, i, j, k, local, evolve, shell, q, h, o)) {
        float* ob = qout + B.off;
#pragma unroll
        for (int c = 0; c < 5; ++c) ob[c * cs + local] = o[c];
    }
}

// ---------------------------------------------------------------------------
// Semi-implicit Helmholtz sweep and strong boundary condition (IMEX path)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int box_of(long long g, const long long* prefix, int nboxes) {
    int b = 0;
    while (b + 1 < nboxes && prefix[b + 1] <= g) ++b;
    return b;
}

// One Jacobi sweep of (I - alpha lap) Q_new = RHS (QSR.cu:786-849) over the
// step's tiles, streamed along i like the L1 stencil (x-register queue, one LDS
// plane with halo 1).  Frozen strong shells and non-evolving voxels keep Q_old;
// box faces are never tiles: face_copy_kernel makes both buffers hold the
// step's input there before the sweeps.  coef[b] = (alpha/dx^2, alpha/dy^2,
// alpha/dz^2, 1 / (1 + 2 alpha sum 1/d^2)).
//
// Converged sweeps are skipped without changing the result.  The sweeps
// ping-pong between two buffers, so the output buffer of sweep s still holds the
// iterate of sweep s-2 (for s = 1: the step's input, the iterate before sweep
// 0).  Sweep s sets flags[row][s] unless every value it writes equals what its
// output buffer held: then out(s) = out(s-2), the FP32 iteration has reached a
// fixed point or a two-cycle (last-bit flips between neighbours), both buffers
// already hold the two states every later sweep would alternate between, and
// those sweeps return at once.  Sweep 0 always counts as changed.  Sweep s also
// clears flags[row ^ 1][s] for the next step (rows alternate between steps).
template <int TK, int TJ, bool BUF>
__global__ __launch_bounds__(TK* TJ) void jacobi_kernel(const float* __restrict__ rhs, const float* __restrict__ qin,
                                                         float* __restrict__ qout, long long cs,
                                                         const BoxDev* __restrict__ boxes,
                                                         const int4* __restrict__ tiles,
                                                         const RowMask* __restrict__ rows, int ch,
                                                         const float4* __restrict__ coef,
                                                         unsigned* __restrict__ flags, int sweep, int nsweeps,
                                                         int row) {
    constexpr int W = TK + 2, R = TJ + 2, PL = W * R, SLOT = 5 * PL;
    __shared__ float lds[2 * SLOT];
    unsigned* fl_cur = flags + row * nsweeps;
    if (nsweeps > 1) {
        if (blockIdx.x == 0 && threadIdx.x == 0) flags[(row ^ 1) * nsweeps + sweep] = 0u;
        if (sweep > 0 && fl_cur[sweep - 1] == 0u) return;  // the same decision in every block
    }
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile (uniform over the block)
    const BoxDev B = boxes[t.x];
    const float4 cf = coef[t.x];
    const int tk = threadIdx.x & (TK - 1), tj = threadIdx.x / TK;
    const int Nx = B.Nx, Ny = B.Ny, Nz = B.Nz;
    const bool per = (B.flags & F_PERIODIC) != 0;
    int ib = t.y, ie = min(t.y + ch, Nx);
    if (!per) {
        ib = max(ib, 1);
        ie = min(ie, Nx - 1);
    }
    if (ib >= ie) return;  // uniform over the block
    const int j = t.z + tj, k = t.w + tk;
    const bool out_ok = per ? (j < Ny && k < Nz) : (j >= 1 && j <= Ny - 2 && k >= 1 && k <= Nz - 2);
    const int jc = wrap_or_clamp(j, Ny, per), kc = wrap_or_clamp(k, Nz, per);
    // halo ring of the plane, one point per thread tid < 2 TK + 2 TJ (as in step_l1_kernel)
    const int tid = threadIdx.x;
    const bool halo = tid < 2 * TK + 2 * TJ;
    int hj = 0, hk = 0;
    if (tid < TK) { hj = -1; hk = tid; }
    else if (tid < 2 * TK) { hj = TJ; hk = tid - TK; }
    else if (tid < 2 * TK + TJ) { hj = tid - 2 * TK; hk = -1; }
    else if (halo) { hj = tid - 2 * TK - TJ; hk = TK; }
    const long long plane = (long long)Ny * Nz;
    const long long o_own = (long long)jc * Nz + kc;
    const long long o_halo = (long long)wrap_or_clamp(t.z + hj, Ny, per) * Nz + wrap_or_clamp(t.w + hk, Nz, per);
    const Field<BUF> F(qin, qout, cs, B.off);    // iterate in, iterate out
    const Field<BUF> Fr(rhs, qout, cs, B.off);   // right-hand side
    const Field<BUF> Fo(qout, qout, cs, B.off);  // what the output buffer holds (iterate two sweeps back)
    const bool detect = nsweeps > 1 && sweep > 0;

    auto ld5 = [&](const Field<BUF>& A, int p, long long o, float v[5]) {
        const long long pb = (long long)p * plane;
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = A.ld(c, pb, o);
    };
    float qm[5], qc[5], qn[5], hc[5] = {}, hn[5] = {};
    float rc[5] = {}, oc[5] = {};
    ld5(F, wrap_or_clamp(ib - 1, Nx, per), o_own, qm);
    ld5(F, ib, o_own, qc);
    ld5(F, wrap_or_clamp(ib + 1, Nx, per), o_own, qn);
    if (halo) ld5(F, ib, o_halo, hc);
    if (out_ok) {
        ld5(Fr, ib, o_own, rc);
        if (detect) ld5(Fo, ib, o_own, oc);
    }
    const int lrow = (tj + 1) * W + tk + 1;
    const int lhalo = (hj + 1) * W + hk + 1;
    const bool pre_rows = out_ok && !per;
    int4 mcur = make_int4(0, 0, 0, 0);
    if (pre_rows) mcur = load_row4(rows, B, ib, j);
    bool changed = false;
    for (int p = ib; p < ie; ++p) {
        // next plane's halo, row mask, RHS and old output, and own Q(p+2): a whole iteration ahead of use
        const bool more = p + 1 < ie;
        const int pn = wrap_or_clamp(p + 1, Nx, per);
        int4 mnext = make_int4(0, 0, 0, 0);
        float qf[5], rn[5] = {}, on[5] = {};
        if (more) {
            if (pre_rows) mnext = load_row4(rows, B, p + 1, j);
            ld5(F, wrap_or_clamp(p + 2, Nx, per), o_own, qf);
            if (halo) ld5(F, pn, o_halo, hn);
            if (out_ok) {
                ld5(Fr, pn, o_own, rn);
                if (detect) ld5(Fo, pn, o_own, on);
            }
        }
        float* L = lds + (p & 1) * SLOT;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            L[c * PL + lrow] = qc[c];
            if (halo) L[c * PL + lhalo] = hc[c];
        }
        lds_barrier();
        if (out_ok) {
            bool evolve, shell;
            decode_row4(mcur, B, k, evolve, shell);
            const bool frozen = !evolve || ((B.flags & F_SKIPSHELL) && shell);
            const long long inpl = (long long)j * Nz + k, pb = (long long)p * plane;
            if (frozen) {
                if (evolve || ((B.flags & F_BC) && shell)) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) F.st(c, pb, inpl, qc[c]);
                }
            } else {
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const float* Lc = L + c * PL;
                    const float nv = (rc[c] + cf.x * (qn[c] + qm[c]) + cf.y * (Lc[lrow + W] + Lc[lrow - W]) +
                                      cf.z * (Lc[lrow + 1] + Lc[lrow - 1])) *
                                     cf.w;
                    if (detect) changed |= __float_as_uint(nv) != __float_as_uint(oc[c]);
                    F.st(c, pb, inpl, nv);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            qm[c] = qc[c];
            qc[c] = qn[c];
            qn[c] = more ? qf[c] : qn[c];
            hc[c] = hn[c];
            rc[c] = rn[c];
            oc[c] = on[c];
        }
        mcur = mnext;
    }
    if (nsweeps > 1) {
        if (sweep == 0) {
            if (blockIdx.x == 0 && threadIdx.x == 0) fl_cur[0] = 1u;
        } else {
            const unsigned long long m = __ballot(changed);
            if (m != 0ull && int(threadIdx.x & 63) == __ffsll((long long)m) - 1) fl_cur[sweep] = 1u;
        }
    }
}

// Box-face voxels that evolve or carry a strong shell: src -> dst, so both
// buffers hold the step's input there (the semi-implicit sweeps keep faces).
__global__ __launch_bounds__(256) void face_copy_kernel(const float* __restrict__ src, float* __restrict__ dst,
                                                        long long cs, const BoxDev* __restrict__ boxes,
                                                        const int4* __restrict__ face_list, long long nlive) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nlive) return;
    const int4 e = face_list[g];
    const BoxDev B = boxes[e.x];
    if (B.flags & (F_PERIODIC | F_INACTIVE)) return;
    const long long v = ((long long)e.y * B.Ny + e.z) * B.Nz + e.w;
#pragma unroll
    for (int c = 0; c < 5; ++c) dst[c * cs + B.off + v] = src[c * cs + B.off + v];
}

// Strong anchoring re-imposed on the shell after the sweeps (QSR.cu:851-867)
__global__ __launch_bounds__(256) void bc_kernel(float* __restrict__ q, long long cs, const BoxDev* __restrict__ boxes,
                                                 const StepDyn* __restrict__ dyn,
                                                 const long long* __restrict__ vox_prefix, int nboxes,
                                                 const RowMask* __restrict__ rows) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= vox_prefix[nboxes]) return;
    const int b = box_of(g, vox_prefix, nboxes);
    const BoxDev B = boxes[b];
    if (!(B.flags & F_BC) || (B.flags & F_PERIODIC)) return;
    const long long v = g - vox_prefix[b];
    const int Nz = B.Nz, Ny = B.Ny;
    const int k = int(v % Nz), j = int((v / Nz) % Ny), i = int(v / ((long long)Ny * Nz));
    bool evolve, shell;
    decode_mask(rows, B, i, j, k, evolve, shell);
    if (!shell) return;
    float q0[5];
    shell_target(B, i, j, k, dyn[b].S_shell, q0);
#pragma unroll
    for (int c = 0; c < 5; ++c) q[c * cs + B.off + v] = q0[c];
}

// ---------------------------------------------------------------------------
// Interchange-format conversion (AoS double <-> SoA float)
// ---------------------------------------------------------------------------
__global__ void aos_to_soa_kernel(const double* __restrict__ aos, float* __restrict__ q0, float* __restrict__ q1,
                                  long long cs, long long off, long long n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        const float x = float(aos[5 * v + c]);
        q0[c * cs + off + v] = x;
        q1[c * cs + off + v] = x;
    }
}

__global__ void soa_to_aos_kernel(const float* __restrict__ q, double* __restrict__ aos, long long cs, long long off,
                                  long long n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) aos[5 * v + c] = double(q[c * cs + off + v]);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace qsrk {

namespace {

template <int V, int EL, bool NZ, bool BF>
void go_aniso(dim3 grid, dim3 block, hipStream_t st, const float* qin, float* qout, long long cs, const BoxDev* boxes,
              const int4* tiles, const StepDyn* dyn, int nboxes, const RowMask* rows, int ch, const float4* imex) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(step_aniso_kernel<tile_k(V), tile_j(V), NZ, BF, EL>), grid, block, 0, st, qin,
                       qout, cs, boxes, tiles, dyn, nboxes, rows, ch, imex);
}

template <int V, bool NZ, bool BF>
void go_l1(dim3 grid, dim3 block, hipStream_t st, const float* qin, float* qout, long long cs, const BoxDev* boxes,
           const int4* tiles, const StepDyn* dyn, int nboxes, const RowMask* rows, int ch, const float4* imex,
           int reverse) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(step_l1_kernel<tile_k(V), tile_j(V), NZ, BF>), grid, block, 0, st, qin, qout, cs,
                       boxes, tiles, dyn, nboxes, rows, ch, imex, reverse);
}

template <int V>
void go_step(int elastic, bool noise, bool buf, dim3 grid, dim3 block, hipStream_t st, const float* qin, float* qout,
             long long cs, const BoxDev* boxes, const int4* tiles, const StepDyn* dyn, int nboxes, const RowMask* rows,
             int ch, const float4* imex, int reverse) {
#define QSR_ARGS grid, block, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch, imex
#define QSR_A(EL) \
    (buf ? (noise ? go_aniso<V, EL, true, true>(QSR_ARGS) : go_aniso<V, EL, false, true>(QSR_ARGS)) \
         : (noise ? go_aniso<V, EL, true, false>(QSR_ARGS) : go_aniso<V, EL, false, false>(QSR_ARGS)))
    switch (elastic) {
        case 1: QSR_A(1); break;
        case 2: QSR_A(2); break;
        case 3: QSR_A(3); break;
        default:
            if (buf) {
                if (noise) go_l1<V, true, true>(QSR_ARGS, reverse);
                else go_l1<V, false, true>(QSR_ARGS, reverse);
            } else {
                if (noise) go_l1<V, true, false>(QSR_ARGS, reverse);
                else go_l1<V, false, false>(QSR_ARGS, reverse);
            }
    }
#undef QSR_A
#undef QSR_ARGS
}

}  // namespace

hipError_t launch_step(int elastic, int variant, bool noise, int ntiles, hipStream_t st, const float* qin, float* qout,
                       long long cs, const BoxDev* boxes, const int4* tiles, const StepDyn* dyn, int nboxes,
                       const RowMask* rows, int ch, const float4* imex, int reverse) {
    const dim3 grid(ntiles), block(TILE_THREADS);
    // 32-bit buffer offsets reach every element when 5 planes x cs floats fit 4 GiB
    const bool buf = cs < (1LL << 32) / 20;
    if (variant == 1)
        go_step<1>(elastic, noise, buf, grid, block, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch, imex, reverse);
    else
        go_step<0>(elastic, noise, buf, grid, block, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch, imex, reverse);
    return hipGetLastError();
}

hipError_t launch_step2(int ntiles, hipStream_t st, const float* qin, float* qout, long long cs, const BoxDev* boxes,
                        const int4* tiles, const StepDyn* dyn1, const StepDyn* dyn2, const RowMask* rows, int ch,
                        int reverse) {
    if (ntiles <= 0) return hipSuccess;
    hipLaunchKernelGGL(step_l1x2_kernel, dim3(ntiles), dim3(X2_TK * X2_TJ), 0, st, qin, qout, cs, boxes, tiles, dyn1,
                       dyn2, rows, ch, reverse);
    return hipGetLastError();
}

hipError_t launch_faces(long long nlive, hipStream_t st, const float* qin, float* qout, long long cs,
                        const BoxDev* boxes, const StepDyn* dyn, const int4* face_list, const RowMask* rows) {
    if (nlive <= 0) return hipSuccess;
    const dim3 grid(unsigned((nlive + 255) / 256)), block(256);
    hipLaunchKernelGGL(step_face_kernel, grid, block, 0, st, qin, qout, cs, boxes, dyn, face_list, nlive, rows);
    return hipGetLastError();
}

hipError_t launch_jacobi(int variant, int ntiles, hipStream_t st, const float* rhs, const float* qold, float* qnew,
                         long long cs, const BoxDev* boxes, const int4* tiles, const RowMask* rows, int ch,
                         const float4* coef, unsigned* flags, int sweep, int nsweeps, int row) {
    const bool buf = cs < (1LL << 32) / 20;
#define QSR_J(V, BF)                                                                                              \
    hipLaunchKernelGGL(HIP_KERNEL_NAME(jacobi_kernel<tile_k(V), tile_j(V), BF>), dim3(ntiles), dim3(TILE_THREADS), 0, \
                       st, rhs, qold, qnew, cs, boxes, tiles, rows, ch, coef, flags, sweep, nsweeps, row)
    if (variant == 1) {
        if (buf) QSR_J(1, true);
        else QSR_J(1, false);
    } else {
        if (buf) QSR_J(0, true);
        else QSR_J(0, false);
    }
#undef QSR_J
    return hipGetLastError();
}
hipError_t launch_face_copy(long long nlive, hipStream_t st, const float* src, float* dst, long long cs,
                            const BoxDev* boxes, const int4* face_list) {
    if (nlive <= 0) return hipSuccess;
    hipLaunchKernelGGL(face_copy_kernel, dim3(unsigned((nlive + 255) / 256)), dim3(256), 0, st, src, dst, cs, boxes,
                       face_list, nlive);
    return hipGetLastError();
}

hipError_t launch_bc(long long nvox, hipStream_t st, float* q, long long cs, const BoxDev* boxes, const StepDyn* dyn,
                     const long long* vox_prefix, int nboxes, const RowMask* rows) {
    hipLaunchKernelGGL(bc_kernel, dim3(unsigned((nvox + 255) / 256)), dim3(256), 0, st, q, cs, boxes, dyn, vox_prefix,
                       nboxes, rows);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(hipStream_t st, const double* aos, float* q0, float* q1, long long cs, long long off,
                             long long n) {
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, aos, q0, q1, cs, off, n);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(hipStream_t st, const float* q, double* aos, long long cs, long long off, long long n) {
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, q, aos, cs, off, n);
    return hipGetLastError();
}

}  // namespace qsrk
