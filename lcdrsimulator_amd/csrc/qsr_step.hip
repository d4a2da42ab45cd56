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
// Tiles: a block owns TJ=4 rows (one wavefront each, 64 consecutive k) of
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

// h = m * [A Q - B (Q^2 - tr(Q^2)/3 I) + C tr(Q^2) Q]   (QSR.cu:399-415)
__device__ __forceinline__ void bulk_field(const float q[5], float cA, float cB, float cC, float h[5]) {
    const float xx = q[0], xy = q[1], xz = q[2], yy = q[3], yz = q[4], zz = -(xx + yy);
    const float q2xx = xx * xx + xy * xy + xz * xz;
    const float q2xy = xx * xy + xy * yy + xz * yz;
    const float q2xz = xx * xz + xy * yz + xz * zz;
    const float q2yy = xy * xy + yy * yy + yz * yz;
    const float q2yz = xy * xz + yy * yz + yz * zz;
    const float tr2 = xx * xx + yy * yy + zz * zz + 2.0f * (xy * xy + xz * xz + yz * yz);
    const float t3 = tr2 * (1.0f / 3.0f);
    const float g = cA + cC * tr2;
    h[0] = g * xx - cB * (q2xx - t3);
    h[1] = g * xy - cB * q2xy;
    h[2] = g * xz - cB * q2xz;
    h[3] = g * yy - cB * (q2yy - t3);
    h[4] = g * yz - cB * q2yz;
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
        const unsigned lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const unsigned lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    }
}

// Five N(0,1) variates keyed by (seed, voxel, iteration): uniforms with 23
// random bits, ((x >> 9) + 0.5) * 2^-23, exact in binary32; Box-Muller pairs.
__device__ __forceinline__ void noise5(unsigned long long seed, unsigned long long voxel, unsigned iter, float g[5]) {
    unsigned a[4] = {unsigned(voxel), unsigned(voxel >> 32), iter, 0u};
    unsigned b[4] = {unsigned(voxel), unsigned(voxel >> 32), iter, 1u};
    philox10(a, unsigned(seed), unsigned(seed >> 32));
    philox10(b, unsigned(seed), unsigned(seed >> 32));
    const unsigned r[6] = {a[0], a[1], a[2], a[3], b[0], b[1]};
    float u[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) u[t] = (float(r[t] >> 9) + 0.5f) * (1.0f / 8388608.0f);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        const float rad = sqrtf(-2.0f * logf(u[2 * p]));
        float s, c;
        sincosf(6.283185307179586f * u[2 * p + 1], &s, &c);
        g[2 * p] = rad * c;
        if (2 * p + 1 < 5) g[2 * p + 1] = rad * s;
    }
}

// Shared tail: anchoring penalty, Euler update, noise, cap, strong BC.
// h is m-scaled and already traceless.  Returns false if the voxel is static.
template <bool NOISE>
__device__ __forceinline__ bool finish_voxel(const BoxDev& B, const StepDyn& D, int i, int j, int k,
                                             long long local, bool evolve, bool shell, const float q[5],
                                             float h[5], float out[5]) {
    const int fl = B.flags;
    const bool update = evolve && !((fl & F_SKIPSHELL) && shell);
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

__device__ __forceinline__ void decode_mask(const RowMask* rows, const BoxDev& B, int i, int j, int k,
                                            bool& evolve, bool& shell) {
    if (B.flags & F_PERIODIC) {
        evolve = true;
        shell = false;
        return;
    }
    const RowMask m = rows[B.rows_off + (long long)i * B.Ny + j];
    evolve = in_iv(k, m.v[0], m.v[1]);
    shell = in_iv(k, m.v[2], m.v[3]) && !in_iv(k, m.v[4], m.v[5]);
}

}  // namespace

// ---------------------------------------------------------------------------
// L1 / kappa kernel: x-register queue + one LDS plane (halo 1), double-buffered.
// ---------------------------------------------------------------------------
template <bool NOISE>
__global__ __launch_bounds__(TK* TJ) void step_l1_kernel(const float* __restrict__ qin, float* __restrict__ qout,
                                                          long long cs, const BoxDev* __restrict__ boxes,
                                                          const int4* __restrict__ tiles,
                                                          const StepDyn* __restrict__ dyn, int nboxes,
                                                          const RowMask* __restrict__ rows, int ch) {
    constexpr int W = TK + 2, R = TJ + 2, PL = W * R, SLOT = 5 * PL;
    __shared__ float lds[2 * SLOT];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD-ordered list (uniform over the block)
    const BoxDev B = boxes[t.x];
    const StepDyn D = dyn[t.x];
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
    const bool rowh = (tj == 0) || (tj == TJ - 1);
    const bool colh = (tk == 0) || (tk == TK - 1);
    const int jh = wrap_or_clamp(tj == 0 ? t.z - 1 : t.z + TJ, Ny, per);
    const int kh = wrap_or_clamp(tk == 0 ? t.w - 1 : t.w + TK, Nz, per);
    const long long plane = (long long)Ny * Nz;
    const long long o_own = (long long)jc * Nz + kc, o_row = (long long)jh * Nz + kc, o_col = (long long)jc * Nz + kh;
    const float* base = qin + B.off;
    float* obase = qout + B.off;

    auto ld5 = [&](int p, long long o, float v[5]) {
        const float* s = base + (long long)p * plane + o;
#pragma unroll
        for (int c = 0; c < 5; ++c) v[c] = s[c * cs];
    };
    float qm[5], qc[5], qn[5], hr[5] = {}, hcol[5] = {}, hrn[5] = {}, hcoln[5] = {};
    ld5(wrap_or_clamp(ib - 1, Nx, per), o_own, qm);
    ld5(ib, o_own, qc);
    if (rowh) ld5(ib, o_row, hr);
    if (colh) ld5(ib, o_col, hcol);
    const int lrow = (tj + 1) * W + tk + 1;
    const int hrow = (tj == 0 ? 0 : (TJ + 1)) * W + tk + 1;
    const int hcolx = (tj + 1) * W + (tk == 0 ? 0 : TK + 1);

    for (int p = ib; p < ie; ++p) {
        const int pn = (p + 1 == Nx) ? 0 : p + 1;  // clamped boxes never reach Nx-1 here
        ld5(pn, o_own, qn);
        if (rowh) ld5(pn, o_row, hrn);
        if (colh) ld5(pn, o_col, hcoln);
        float* L = lds + (p & 1) * SLOT;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            L[c * PL + lrow] = qc[c];
            if (rowh) L[c * PL + hrow] = hr[c];
            if (colh) L[c * PL + hcolx] = hcol[c];
        }
        __syncthreads();
        if (out_ok) {
            bool evolve, shell;
            decode_mask(rows, B, p, j, k, evolve, shell);
            const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
            if (live) {
                float h[5];
                bulk_field(qc, D.cA, B.cB, B.cC, h);
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const float* Lc = L + c * PL;
                    const float q2 = 2.0f * qc[c];
                    const float lap = B.lx * (qn[c] + qm[c] - q2) + B.ly * (Lc[lrow + W] + Lc[lrow - W] - q2) +
                                      B.lz * (Lc[lrow + 1] + Lc[lrow - 1] - q2);
                    h[c] -= lap;
                }
                float o[5];
                const long long local = (long long)p * plane + (long long)j * Nz + k;
                if (finish_voxel<NOISE>(B, D, p, j, k, local, evolve, shell, qc, h, o)) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) __builtin_nontemporal_store(o[c], obase + c * cs + local);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            qm[c] = qc[c];
            qc[c] = qn[c];
            hr[c] = hrn[c];
            hcol[c] = hcoln[c];
        }
    }
}

// ---------------------------------------------------------------------------
// Anisotropic kernel (L2 and/or L3): five LDS Q-plane slots with halo 2 and two
// Dcol-plane slots with halo 1; Dcol = div of Q's columns (QSR.cu:325-371) is
// computed once per point on the halo ring and reused by the L2 derivatives
// (QSR.cu:436-460) and the L3 terms (QSR.cu:474-621).
// ---------------------------------------------------------------------------
template <bool NOISE>
__global__ __launch_bounds__(TK* TJ) void step_full_kernel(const float* __restrict__ qin, float* __restrict__ qout,
                                                            long long cs, const BoxDev* __restrict__ boxes,
                                                            const int4* __restrict__ tiles,
                                                            const StepDyn* __restrict__ dyn, int nboxes,
                                                            const RowMask* __restrict__ rows, int ch) {
    constexpr int W2 = TK + 4, R2 = TJ + 4, P2 = W2 * R2, QSLOT = 5 * P2;
    constexpr int W1 = TK + 2, R1 = TJ + 2, P1 = W1 * R1, DSLOT = 3 * P1;
    constexpr int NT = TK * TJ;
    constexpr int NPF = (P2 + NT - 1) / NT;  // plane points per thread
    __shared__ float qs[5 * QSLOT];
    __shared__ float ds[2 * DSLOT];
    const int4 t = tiles[blockIdx.x];
    if (t.x < 0) return;  // padding tile of the XCD-ordered list (uniform over the block)
    const BoxDev B = boxes[t.x];
    const StepDyn D = dyn[t.x];
    const int tid = threadIdx.x, tk = tid & (TK - 1), tj = tid / TK;
    const int Nx = B.Nx, Ny = B.Ny, Nz = B.Nz;
    const bool per = (B.flags & F_PERIODIC) != 0;
    int ib = t.y, ie = min(t.y + ch, Nx);
    if (!per) {
        ib = max(ib, 1);
        ie = min(ie, Nx - 1);
    }
    if (ib >= ie) return;
    const int j0 = t.z, k0 = t.w;
    const int j = j0 + tj, k = k0 + tk;
    const bool out_ok = per ? (j < Ny && k < Nz) : (j >= 1 && j <= Ny - 2 && k >= 1 && k <= Nz - 2);
    const long long plane = (long long)Ny * Nz;
    const float* base = qin + B.off;
    float* obase = qout + B.off;
    const bool hasL2 = (B.flags & F_L2) != 0, hasL3 = (B.flags & F_L3) != 0;

    // per-thread in-plane offsets of the plane points it loads
    long long poff[NPF];
    int pidx[NPF];
#pragma unroll
    for (int s = 0; s < NPF; ++s) {
        const int id = tid + s * NT;
        pidx[s] = id < P2 ? id : -1;
        const int r = id / W2, c = id % W2;
        const int jj = wrap_or_clamp(j0 - 2 + r, Ny, per), kk = wrap_or_clamp(k0 - 2 + c, Nz, per);
        poff[s] = (long long)jj * Nz + kk;
    }
    float pf[NPF][5];
    auto fetch = [&](int p) {
        const long long pb = (long long)wrap_or_clamp(p, Nx, per) * plane;
#pragma unroll
        for (int s = 0; s < NPF; ++s)
            if (pidx[s] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) pf[s][c] = base[c * cs + pb + poff[s]];
            }
    };
    auto commit = [&](int slot) {
        float* Q = qs + slot * QSLOT;
#pragma unroll
        for (int s = 0; s < NPF; ++s)
            if (pidx[s] >= 0) {
#pragma unroll
                for (int c = 0; c < 5; ++c) Q[c * P2 + pidx[s]] = pf[s][c];
            }
    };
    // first-derivative scale along an axis at coordinate x: one-sided on faces
    auto dscale = [&](int x, int n, float half, float full) {
        return (!per && (x == 0 || x == n - 1)) ? full : half;
    };
    // Dcol on the (TJ+2) x (TK+2) ring-including region of plane p (loop plane
    // index, may be outside [0, Nx) only for clamped-away planes that are unused)
    auto dcol = [&](int sm, int sc, int sp, int dslot, int p) {
        const float* Qm = qs + sm * QSLOT;
        const float* Qc = qs + sc * QSLOT;
        const float* Qp = qs + sp * QSLOT;
        float* Dd = ds + dslot * DSLOT;
        const float ax = dscale(p, Nx, B.hx, B.fx);
        for (int id = tid; id < P1; id += NT) {
            const int rr = id / W1, cc = id % W1;
            const int r = rr + 1, c = cc + 1, at = r * W2 + c;
            const int jj = j0 - 1 + rr, kk = k0 - 1 + cc;
            const float ay = dscale(jj, Ny, B.hy, B.fy), az = dscale(kk, Nz, B.hz, B.fz);
            float dx_[5], dy_[5], dz_[5];
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                dx_[q] = (Qp[q * P2 + at] - Qm[q * P2 + at]) * ax;
                dy_[q] = (Qc[q * P2 + at + W2] - Qc[q * P2 + at - W2]) * ay;
                dz_[q] = (Qc[q * P2 + at + 1] - Qc[q * P2 + at - 1]) * az;
            }
            Dd[0 * P1 + id] = dx_[0] + dy_[1] + dz_[2];
            Dd[1 * P1 + id] = dx_[1] + dy_[3] + dz_[4];
            Dd[2 * P1 + id] = dx_[2] + dy_[4] - (dz_[0] + dz_[3]);
        }
    };

    // prologue: planes ib-2 .. ib+1 into slots 0..3 (slot of plane p = (p - ib + 2) % 5)
    auto qslot = [&](int p) { return (p - ib + 2) % 5; };
    auto dslt = [&](int p) { return (p - ib + 2) & 1; };
    for (int p = ib - 2; p <= ib + 1; ++p) {
        fetch(p);
        commit(qslot(p));
    }
    __syncthreads();
    dcol(qslot(ib - 2), qslot(ib - 1), qslot(ib), dslt(ib - 1), ib - 1);
    dcol(qslot(ib - 1), qslot(ib), qslot(ib + 1), dslt(ib), ib);
    __syncthreads();
    const int lq = (tj + 2) * W2 + tk + 2;  // own point in a Q slot
    const int ld = (tj + 1) * W1 + tk + 1;  // own point in a D slot
    float dm[3];
    {
        const float* Dd = ds + dslt(ib - 1) * DSLOT;
#pragma unroll
        for (int a = 0; a < 3; ++a) dm[a] = Dd[a * P1 + ld];
    }
    fetch(ib + 2);

    for (int p = ib; p < ie; ++p) {
        commit(qslot(p + 2));
        __syncthreads();
        dcol(qslot(p), qslot(p + 1), qslot(p + 2), dslt(p + 1), p + 1);
        __syncthreads();
        const float* Qm = qs + qslot(p - 1) * QSLOT;
        const float* Qc = qs + qslot(p) * QSLOT;
        const float* Qp = qs + qslot(p + 1) * QSLOT;
        const float* Dc = ds + dslt(p) * DSLOT;
        const float* Dn = ds + dslt(p + 1) * DSLOT;
        float dcen[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) dcen[a] = Dc[a * P1 + ld];
        if (out_ok) {
            bool evolve, shell;
            decode_mask(rows, B, p, j, k, evolve, shell);
            const bool live = (evolve && !((B.flags & F_SKIPSHELL) && shell)) || ((B.flags & F_BC) && shell);
            if (live) {
                float q[5], sx_[5], sy_[5], sz_[5];
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    const float* A = Qc + c * P2;
                    q[c] = A[lq];
                    sx_[c] = Qp[c * P2 + lq] + Qm[c * P2 + lq] - 2.0f * q[c];
                    sy_[c] = A[lq + W2] + A[lq - W2] - 2.0f * q[c];
                    sz_[c] = A[lq + 1] + A[lq - 1] - 2.0f * q[c];
                }
                float h[5];
                bulk_field(q, D.cA, B.cB, B.cC, h);
#pragma unroll
                for (int c = 0; c < 5; ++c) h[c] -= B.lx * sx_[c] + B.ly * sy_[c] + B.lz * sz_[c];
                if (hasL2) {
                    // central derivatives of Dcol at an interior voxel
                    float dn[3], dyp[3], dym[3], dzp[3], dzm[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        dn[a] = Dn[a * P1 + ld];
                        dyp[a] = Dc[a * P1 + ld + W1];
                        dym[a] = Dc[a * P1 + ld - W1];
                        dzp[a] = Dc[a * P1 + ld + 1];
                        dzm[a] = Dc[a * P1 + ld - 1];
                    }
                    const float dDx_dx = (dn[0] - dm[0]) * B.hx, dDy_dx = (dn[1] - dm[1]) * B.hx, dDz_dx = (dn[2] - dm[2]) * B.hx;
                    const float dDx_dy = (dyp[0] - dym[0]) * B.hy, dDy_dy = (dyp[1] - dym[1]) * B.hy, dDz_dy = (dyp[2] - dym[2]) * B.hy;
                    const float dDx_dz = (dzp[0] - dzm[0]) * B.hz, dDy_dz = (dzp[1] - dzm[1]) * B.hz, dDz_dz = (dzp[2] - dzm[2]) * B.hz;
                    const float gxx = B.mL2 * dDx_dx, gyy = B.mL2 * dDy_dy, gzz = B.mL2 * dDz_dz;
                    const float tr = (gxx + gyy + gzz) * (1.0f / 3.0f);
                    h[0] -= gxx - tr;
                    h[3] -= gyy - tr;
                    h[1] -= B.mL2 * (dDy_dx + dDx_dy);
                    h[2] -= B.mL2 * (dDz_dx + dDx_dz);
                    h[4] -= B.mL2 * (dDz_dy + dDy_dz);
                }
                if (hasL3) {
                    // first derivatives d[c][axis], c = xx,xy,xz,yy,yz,zz
                    float d[6][3];
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const float* A = Qc + c * P2;
                        d[c][0] = (Qp[c * P2 + lq] - Qm[c * P2 + lq]) * B.hx;
                        d[c][1] = (A[lq + W2] - A[lq - W2]) * B.hy;
                        d[c][2] = (A[lq + 1] - A[lq - 1]) * B.hz;
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
                    float M[3][3];
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = a; b < 3; ++b)
                            M[a][b] = d[0][a] * d[0][b] + d[3][a] * d[3][b] + d[5][a] * d[5][b] +
                                      2.0f * (d[1][a] * d[1][b] + d[2][a] * d[2][b] + d[4][a] * d[4][b]);
                    const float xx = q[0], xy = q[1], xz = q[2], yy = q[3], yz = q[4], zz = -(xx + yy);
                    float Bc[5];
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        const float* A = Qc + c * P2;
                        const float* Ap = Qp + c * P2;
                        const float* Am = Qm + c * P2;
                        const float d2xx = sx_[c] * B.ix2, d2yy = sy_[c] * B.iy2, d2zz = sz_[c] * B.iz2;
                        const float d2xy = (Ap[lq + W2] - Ap[lq - W2] - Am[lq + W2] + Am[lq - W2]) * B.i4xy;
                        const float d2xz = (Ap[lq + 1] - Ap[lq - 1] - Am[lq + 1] + Am[lq - 1]) * B.i4xz;
                        const float d2yz = (A[lq + W2 + 1] - A[lq + W2 - 1] - A[lq - W2 + 1] + A[lq - W2 - 1]) * B.i4yz;
                        const float contract = xx * d2xx + yy * d2yy + zz * d2zz + 2.0f * (xy * d2xy + xz * d2xz + yz * d2yz);
                        Bc[c] = dcen[0] * d[c][0] + dcen[1] * d[c][1] + dcen[2] * d[c][2] + contract;
                    }
                    // h += (L3/2) M - L3 B, B_zz = -(B_xx + B_yy), then remove the trace
                    const float hl = 0.5f * B.mL3;
                    const float tr = hl * (M[0][0] + M[1][1] + M[2][2]) * (1.0f / 3.0f);
                    h[0] += hl * M[0][0] - B.mL3 * Bc[0] - tr;
                    h[3] += hl * M[1][1] - B.mL3 * Bc[3] - tr;
                    h[1] += hl * M[0][1] - B.mL3 * Bc[1];
                    h[2] += hl * M[0][2] - B.mL3 * Bc[2];
                    h[4] += hl * M[1][2] - B.mL3 * Bc[4];
                }
                float o[5];
                const long long local = (long long)p * plane + (long long)j * Nz + k;
                if (finish_voxel<NOISE>(B, D, p, j, k, local, evolve, shell, q, h, o)) {
#pragma unroll
                    for (int c = 0; c < 5; ++c) __builtin_nontemporal_store(o[c], obase + c * cs + local);
                }
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) dm[a] = dcen[a];
        if (p + 1 < ie) fetch(p + 3);
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

__device__ __forceinline__ void face_decode(long long f, int Nx, int Ny, int Nz, int& i, int& j, int& k) {
    const long long a = (long long)Ny * Nz, b = (long long)(Nx - 2) * Nz, c = (long long)(Nx - 2) * (Ny - 2);
    if (f < 2 * a) {
        i = f < a ? 0 : Nx - 1;
        const long long r = f % a;
        j = int(r / Nz);
        k = int(r % Nz);
        return;
    }
    f -= 2 * a;
    if (f < 2 * b) {
        j = f < b ? 0 : Ny - 1;
        const long long r = f % b;
        i = 1 + int(r / Nz);
        k = int(r % Nz);
        return;
    }
    f -= 2 * b;
    k = f < c ? 0 : Nz - 1;
    const long long r = f % c;
    i = 1 + int(r / (Ny - 2));
    j = 1 + int(r % (Ny - 2));
}

}  // namespace

__global__ __launch_bounds__(256) void step_face_kernel(const float* __restrict__ qin, float* __restrict__ qout,
                                                        long long cs, const BoxDev* __restrict__ boxes,
                                                        const StepDyn* __restrict__ dyn,
                                                        const long long* __restrict__ face_prefix, int nboxes,
                                                        const RowMask* __restrict__ rows) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= face_prefix[nboxes]) return;
    int b = 0;
    while (b + 1 < nboxes && face_prefix[b + 1] <= g) ++b;
    const BoxDev B = boxes[b];
    if (B.flags & F_PERIODIC) return;
    const StepDyn D = dyn[b];
    int i, j, k;
    face_decode(g - face_prefix[b], B.Nx, B.Ny, B.Nz, i, j, k);
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

// Explicit instantiations used by the context code.
template __global__ void step_l1_kernel<false>(const float*, float*, long long, const BoxDev*, const int4*,
                                               const StepDyn*, int, const RowMask*, int);
template __global__ void step_l1_kernel<true>(const float*, float*, long long, const BoxDev*, const int4*,
                                              const StepDyn*, int, const RowMask*, int);
template __global__ void step_full_kernel<false>(const float*, float*, long long, const BoxDev*, const int4*,
                                                 const StepDyn*, int, const RowMask*, int);
template __global__ void step_full_kernel<true>(const float*, float*, long long, const BoxDev*, const int4*,
                                                const StepDyn*, int, const RowMask*, int);

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace qsrk {

hipError_t launch_step(bool full, bool noise, int ntiles, hipStream_t st, const float* qin, float* qout,
                       long long cs, const BoxDev* boxes, const int4* tiles, const StepDyn* dyn, int nboxes,
                       const RowMask* rows, int ch) {
    const dim3 grid(ntiles), block(TK * TJ);
    if (full) {
        if (noise) hipLaunchKernelGGL(step_full_kernel<true>, grid, block, 0, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch);
        else hipLaunchKernelGGL(step_full_kernel<false>, grid, block, 0, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch);
    } else {
        if (noise) hipLaunchKernelGGL(step_l1_kernel<true>, grid, block, 0, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch);
        else hipLaunchKernelGGL(step_l1_kernel<false>, grid, block, 0, st, qin, qout, cs, boxes, tiles, dyn, nboxes, rows, ch);
    }
    return hipGetLastError();
}

hipError_t launch_faces(long long nface, hipStream_t st, const float* qin, float* qout, long long cs,
                        const BoxDev* boxes, const StepDyn* dyn, const long long* face_prefix, int nboxes,
                        const RowMask* rows) {
    if (nface <= 0) return hipSuccess;
    const dim3 grid(unsigned((nface + 255) / 256)), block(256);
    hipLaunchKernelGGL(step_face_kernel, grid, block, 0, st, qin, qout, cs, boxes, dyn, face_prefix, nboxes, rows);
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
