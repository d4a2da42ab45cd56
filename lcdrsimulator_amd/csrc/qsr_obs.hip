// qsr_obs.hip -- FP64 log-point observables on the GPU.
//
// Replaces the reference's log pass, which copies whole per-voxel arrays to
// the host and sums them there (QSR.cu:2323-2341, 2788-2822, 2827-2874;
// KZM_bulk_ldg.cu:923-925, 543-713).  Per-voxel quantities are computed in FP64
// from the FP32 field; FP contraction is off so the FP64 arithmetic (nematic
// power iteration, geometry, energy density) rounds exactly like the host
// restatement.  Block partial sums use DPP wave reductions and are combined in
// a fixed order, so results are run-to-run deterministic.
#pragma clang fp contract(off)
#include <algorithm>

#include "qsr_kernels.hpp"

using namespace qsrk;


using namespace qsro;

namespace {

// ---- DPP wave reductions (result valid in every lane) ----------------------
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
// quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E, row_half_mirror = 0x141,
// row_mirror = 0x140: after the four steps every lane holds its 16-lane row total.
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_f64<0xB1>(v);
    v = v + dpp_f64<0x4E>(v);
    v = v + dpp_f64<0x141>(v);
    v = v + dpp_f64<0x140>(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}
__device__ __forceinline__ double wave_max(double v) {
    v = fmax(v, dpp_f64<0xB1>(v));
    v = fmax(v, dpp_f64<0x4E>(v));
    v = fmax(v, dpp_f64<0x141>(v));
    v = fmax(v, dpp_f64<0x140>(v));
    return fmax(fmax(readlane_f64(v, 0), readlane_f64(v, 16)), fmax(readlane_f64(v, 32), readlane_f64(v, 48)));
}

// Block (256 threads = 4 waves) reduction of N slots; `is_max` marks max slots.
template <int N>
__device__ __forceinline__ void block_reduce_store(double (&v)[N], const bool (&is_max)[N], double* dst) {
    __shared__ double red[4][N];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int s = 0; s < N; ++s) {
        const double r = is_max[s] ? wave_max(v[s]) : wave_sum(v[s]);
        if (lane == 0) red[wave][s] = r;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        const int s = threadIdx.x;
        double r = red[0][s];
        for (int w = 1; w < 4; ++w) r = is_max[s] ? fmax(r, red[w][s]) : r + red[w][s];
        dst[s] = r;
    }
}

__device__ __forceinline__ bool in_iv(int k, short lo, short hi) { return k >= lo && k <= hi; }

// calculateNematicFieldDevice (QSR.cu:191-243); kzm selects the periodic
// branch's variant (KZM_bulk_ldg.cu:200-252: threshold 1e-12, n = (1,0,0)).
__device__ void nematic(const double q5[5], bool kzm, double out[4]) {
    const double xx = q5[0], xy = q5[1], xz = q5[2], yy = q5[3], yz = q5[4], zz = -xx - yy;
    const double thr = kzm ? 1e-12 : 1e-9;
    if (fabs(xx) < thr && fabs(xy) < thr && fabs(xz) < thr && fabs(yy) < thr && fabs(yz) < thr) {
        out[0] = 0.0;
        out[1] = kzm ? 1.0 : 0.0;
        out[2] = 0.0;
        out[3] = kzm ? 0.0 : 1.0;
        return;
    }
    double nx = xx + xy + xz, ny = xy + yy + yz, nz = xz + yz - xx - yy;
    const double n0 = sqrt(nx * nx + ny * ny + nz * nz);
    if (n0 > 1e-12) {
        const double inv = 1.0 / n0;  // one divide, three products
        nx *= inv; ny *= inv; nz *= inv;
    } else {
        nx = 0.0; ny = 0.0; nz = 1.0;
    }
    const double r1 = fabs(xx) + fabs(xy) + fabs(xz);
    const double r2 = fabs(xy) + fabs(yy) + fabs(yz);
    const double r3 = fabs(xz) + fabs(yz) + fabs(zz);
    const double alpha = fmax(r1, fmax(r2, r3)) + 1e-12;
    for (int it = 0; it < 15; ++it) {
        const double ax = xx * nx + xy * ny + xz * nz + alpha * nx;
        const double ay = xy * nx + yy * ny + yz * nz + alpha * ny;
        const double az = xz * nx + yz * ny + zz * nz + alpha * nz;
        const double nrm = sqrt(ax * ax + ay * ay + az * az);
        if (nrm < 1e-13) break;
        const double inv = 1.0 / nrm;
        nx = ax * inv; ny = ay * inv; nz = az * inv;
    }
    const double lam = nx * (xx * nx + xy * ny + xz * nz) + ny * (xy * nx + yy * ny + yz * nz) +
                       nz * (xz * nx + yz * ny + zz * nz);
    if (nz < -1e-10) {
        nx = -nx; ny = -ny; nz = -nz;
    }
    out[0] = 1.5 * lam; out[1] = nx; out[2] = ny; out[3] = nz;
}

__device__ __forceinline__ void load_q(const float* q, long long cs, long long v, double out[5]) {
#pragma unroll
    for (int c = 0; c < 5; ++c) out[c] = double(q[c * cs + v]);
}

// wrap_to_pi_device (QSR.cu:1027-1035)
__device__ __forceinline__ double wrap_floor(double x) {
    const double kPi = 3.141592653589793238462643383279502884, kTwoPi = 2.0 * kPi;
    x = x + kPi;
    x -= floor(x / kTwoPi) * kTwoPi;
    return x - kPi;
}
// wrap_to_pi (KZM_bulk_ldg.cu:471-479)
__device__ __forceinline__ double wrap_loop(double d) {
    const double kPi = 3.14159265358979323846;
    while (d <= -kPi) d += 2.0 * kPi;
    while (d > kPi) d -= 2.0 * kPi;
    return d;
}

}  // namespace

// ---------------------------------------------------------------------------
// Volume pass, droplet box: energy (QSR.cu:869-992), anchoring energy
// (QSR.cu:654-695), avg/max S (QSR.cu:2285-2321), radiality (QSR.cu:994-1015).
// partial slots: bulk, elastic, anchoring, S_sum, S_cnt, rad_sum, rad_cnt, max_S
// ---------------------------------------------------------------------------
// Block bid of an nblk-block droplet volume pass over one box: its partials (NPART slots) into
// partials + bid * NPART.
__device__ __forceinline__ void droplet_block(const float* __restrict__ q, long long cs, const ObsParams& P,
                                              const RowMask* __restrict__ step_rows,
                                              const RowMask* __restrict__ obs_rows, double* __restrict__ partials,
                                              long long bid, long long nblk) {
    double acc[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    // rows (i, j) in a grid stride, a thread's k fixed within a row: each thread accumulates its
    // voxels in a fixed order (deterministic sums) and splits a row index once, not per voxel
    const long long rows = (long long)P.Nx * P.Ny;
    const int kw = P.Nz < 256 ? P.Nz : 256;  // threads per row
    const int rpi = 256 / kw;                 // rows per block iteration
    const int rsub = threadIdx.x / kw, k0 = threadIdx.x - rsub * kw;
    for (long long r = bid * rpi + rsub; rsub < rpi && r < rows; r += nblk * rpi) {
      const int i = int(r / P.Ny), j = int(r - (long long)i * P.Ny);
      for (int k = k0; k < P.Nz; k += kw) {
        const long long v = r * P.Nz + k;
        const float* base = q + P.off;
        double Q[5];
        load_q(base, cs, v, Q);
        const double dV = P.dx * P.dy * P.dz;
        const double xx = Q[0], xy = Q[1], xz = Q[2], yy = Q[3], yz = Q[4], zz = -xx - yy;
        if (i >= 1 && i < P.Nx - 1 && j >= 1 && j < P.Ny - 1 && k >= 1 && k < P.Nz - 1) {
            const double tr2 = xx * xx + yy * yy + zz * zz + 2 * (xy * xy + xz * xz + yz * yz);
            const double q2xx = xx * xx + xy * xy + xz * xz, q2xy = xx * xy + xy * yy + xz * yz;
            const double q2xz = xx * xz + xy * yz + xz * zz, q2yx = xy * xx + yy * xy + yz * xz;
            const double q2yy = xy * xy + yy * yy + yz * yz, q2yz = xy * xz + yy * yz + yz * zz;
            const double q2zx = xz * xx + yz * xy + zz * xz, q2zy = xz * xy + yz * yy + zz * yz;
            const double q2zz = xz * xz + yz * yz + zz * zz;
            const double tr3 = xx * q2xx + xy * q2yx + xz * q2zx + xy * q2xy + yy * q2yy + yz * q2zy + xz * q2xz +
                               yz * q2yz + zz * q2zz;
            acc[0] += (P.c2 * tr2 - P.c3 * tr3 + P.c4 * tr2 * tr2) * dV;
            const long long sx = (long long)P.Ny * P.Nz, sy = P.Nz;
            const double i2x = 1.0 / (2.0 * P.dx), i2y = 1.0 / (2.0 * P.dy), i2z = 1.0 / (2.0 * P.dz);
            double d[6][3];
            for (int c = 0; c < 5; ++c) {
                const float* b = base + c * cs;
                d[c][0] = (double(b[v + sx]) - double(b[v - sx])) * i2x;
                d[c][1] = (double(b[v + sy]) - double(b[v - sy])) * i2y;
                d[c][2] = (double(b[v + 1]) - double(b[v - 1])) * i2z;
            }
            for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
            auto sq3 = [](const double* w) { return w[0] * w[0] + w[1] * w[1] + w[2] * w[2]; };
            const double g2 = sq3(d[0]) + sq3(d[3]) + sq3(d[5]) + 2.0 * (sq3(d[1]) + sq3(d[2]) + sq3(d[4]));
            double fel;
            if (P.no_L) {
                fel = 0.5 * P.kappa * g2;
            } else {
                const double f1 = 0.5 * P.L1 * g2;
                const double dvx = d[0][0] + d[1][1] + d[2][2];
                const double dvy = d[1][0] + d[3][1] + d[4][2];
                const double dvz = d[2][0] + d[4][1] + d[5][2];
                const double f2 = 0.5 * P.L2 * (dvx * dvx + dvy * dvy + dvz * dvz);
                double f3 = 0.0;
                if (P.L3 != 0.0) {
                    double M[3][3];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b)
                            M[a][b] = d[0][a] * d[0][b] + d[3][a] * d[3][b] + d[5][a] * d[5][b] +
                                      2.0 * (d[1][a] * d[1][b] + d[2][a] * d[2][b] + d[4][a] * d[4][b]);
                    f3 = 0.5 * P.L3 * (xx * M[0][0] + yy * M[1][1] + zz * M[2][2] +
                                       2.0 * (xy * M[0][1] + xz * M[0][2] + yz * M[1][2]));
                }
                fel = f1 + f2 + f3;
            }
            acc[1] += fel * dV;
        }
        const RowMask sm = step_rows[P.rows_off + (long long)i * P.Ny + j];
        const RowMask om = obs_rows[P.rows_off + (long long)i * P.Ny + j];
        const bool shell = in_iv(k, sm.v[2], sm.v[3]) && !in_iv(k, sm.v[4], sm.v[5]);
        const double x = i * P.dx - P.cx, y = j * P.dy - P.cy, z = k * P.dz - P.cz;
        if (shell && P.W_over_delta != 0.0) {
            // getSurfaceNormal (QSR.cu:130-142)
            const double gx = x / (P.rx * P.rx), gy = y / (P.ry * P.ry), gz = z / (P.rz * P.rz);
            const double gn = sqrt(gx * gx + gy * gy + gz * gz);
            double nx, ny, nz;
            if (gn <= 1e-30) {
                const double r = sqrt(x * x + y * y + z * z);
                if (r <= 1e-30) { nx = 1.0; ny = 0.0; nz = 0.0; }
                else { nx = x / r; ny = y / r; nz = z / r; }
            } else {
                nx = gx / gn; ny = gy / gn; nz = gz / gn;
            }
            const double S = P.S_shell;
            const double q0xx = S * (nx * nx - 1.0 / 3.0), q0xy = S * (nx * ny), q0xz = S * (nx * nz);
            const double q0yy = S * (ny * ny - 1.0 / 3.0), q0yz = S * (ny * nz), q0zz = -(q0xx + q0yy);
            const double a = xx - q0xx, b = xy - q0xy, c = xz - q0xz, e = yy - q0yy, f = yz - q0yz;
            const double zd = (-(xx + yy)) - q0zz;
            const double diff2 = a * a + e * e + zd * zd + 2.0 * (b * b + c * c + f * f);
            acc[2] += 0.5 * P.W_over_delta * diff2 * dV;
        }
        const bool inA = in_iv(k, om.v[0], om.v[1]);
        const bool inRad = in_iv(k, om.v[2], om.v[3]) && !in_iv(k, om.v[4], om.v[5]);
        if (!P.skip_order && (inA || inRad)) {
            double nf[4];
            nematic(Q, false, nf);
            if (inA) {
                acc[3] += nf[0];
                acc[4] += 1.0;
                if (isfinite(nf[0]) && nf[0] > acc[7]) acc[7] = nf[0];
            }
            if (inRad) {
                const double r = sqrt(x * x + y * y + z * z);
                double rx = 1.0, ry = 0.0, rz = 0.0;
                if (r > 1e-30) { rx = x / r; ry = y / r; rz = z / r; }
                acc[5] += fabs(nf[1] * rx + nf[2] * ry + nf[3] * rz);
                acc[6] += 1.0;
            }
        }
      }
    }
    const bool is_max[NPART] = {false, false, false, false, false, false, false, true};
    block_reduce_store<NPART>(acc, is_max, partials + bid * NPART);
}

__global__ __launch_bounds__(256) void obs_droplet_kernel(const float* __restrict__ q, long long cs, ObsParams P,
                                                          const RowMask* __restrict__ step_rows,
                                                          const RowMask* __restrict__ obs_rows,
                                                          double* __restrict__ partials) {
    droplet_block(q, cs, P, step_rows, obs_rows, partials, blockIdx.x, gridDim.x);
}

// The droplet volume pass of n boxes in one launch: global block g belongs to box m with
// prefix[m] <= g < prefix[m+1] and computes exactly block g - prefix[m] of obs_droplet_kernel's
// grid for that box (prefix[m+1] - prefix[m] blocks), writing its partials to partials + g * NPART.
__global__ __launch_bounds__(256) void obs_droplet_multi_kernel(const float* __restrict__ q, long long cs,
                                                                const ObsParams* __restrict__ Ps,
                                                                const long long* __restrict__ prefix, int n,
                                                                const RowMask* __restrict__ step_rows,
                                                                const RowMask* __restrict__ obs_rows,
                                                                double* __restrict__ partials) {
    const long long g = blockIdx.x;
    int lo = 0, hi = n - 1;  // the box whose block range holds g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    const ObsParams P = Ps[lo];
    droplet_block(q, cs, P, step_rows, obs_rows, partials + prefix[lo] * NPART, g - prefix[lo],
                  prefix[lo + 1] - prefix[lo]);
}

// ---------------------------------------------------------------------------
// Volume pass, periodic box (KZM_bulk_ldg.cu:567-633): forward-difference
// energies, integrals of trQ^2 and |grad Q|^2, S stats; stores the nematic
// field for the plaquette pass.
// partial slots: bulk, elastic, tr2_int, grad_int, S_sum, S_cnt, -, max_S
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void obs_periodic_kernel(const float* __restrict__ q, long long cs, ObsParams P,
                                                           double* __restrict__ nf_out, double* __restrict__ partials) {
    const long long n = (long long)P.Nx * P.Ny * P.Nz;
    double acc[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int k = int(v % P.Nz), j = int((v / P.Nz) % P.Ny), i = int(v / ((long long)P.Ny * P.Nz));
        const float* base = q + P.off;
        double Q[5];
        load_q(base, cs, v, Q);
        const long long vx = ((long long)((i + 1) % P.Nx) * P.Ny + j) * P.Nz + k;
        const long long vy = ((long long)i * P.Ny + (j + 1) % P.Ny) * P.Nz + k;
        const long long vz = ((long long)i * P.Ny + j) * P.Nz + (k + 1) % P.Nz;
        const double xx = Q[0], xy = Q[1], xz = Q[2], yy = Q[3], yz = Q[4], zz = -(xx + yy);
        const double q2xx = xx * xx + xy * xy + xz * xz, q2xy = xx * xy + xy * yy + xz * yz;
        const double q2xz = xx * xz + xy * yz + xz * zz, q2yy = xy * xy + yy * yy + yz * yz;
        const double q2yz = xy * xz + yy * yz + yz * zz, q2zz = xz * xz + yz * yz + zz * zz;
        const double tr2 = xx * xx + yy * yy + zz * zz + 2.0 * (xy * xy + xz * xz + yz * yz);
        const double tr3 = xx * q2xx + xy * q2xy + xz * q2xz + xy * q2xy + yy * q2yy + yz * q2yz + xz * q2xz +
                           yz * q2yz + zz * q2zz;
        const double dV = P.dx * P.dy * P.dz;
        acc[0] += (P.c2 * tr2 - P.c3 * tr3 + P.c4 * tr2 * tr2) * dV;
        double d[6][3];
        for (int c = 0; c < 5; ++c) {
            const float* b = base + c * cs;
            d[c][0] = (double(b[vx]) - Q[c]) * (1.0 / P.dx);
            d[c][1] = (double(b[vy]) - Q[c]) * (1.0 / P.dy);
            d[c][2] = (double(b[vz]) - Q[c]) * (1.0 / P.dz);
        }
        for (int a = 0; a < 3; ++a) d[5][a] = -(d[0][a] + d[3][a]);
        double g2 = 0.0, g2o = 0.0;
        for (int a = 0; a < 3; ++a) g2 += d[0][a] * d[0][a] + d[3][a] * d[3][a] + d[5][a] * d[5][a];
        for (int a = 0; a < 3; ++a) g2o += d[1][a] * d[1][a] + d[2][a] * d[2][a] + d[4][a] * d[4][a];
        g2 += 2.0 * g2o;
        acc[1] += 0.5 * P.kappa * g2 * dV;
        acc[2] += tr2 * dV;
        acc[3] += g2 * dV;
        double nf[4];
        nematic(Q, true, nf);
        for (int c = 0; c < 4; ++c) nf_out[4 * v + c] = nf[c];
        acc[4] += nf[0];
        acc[5] += 1.0;
        if (nf[0] > acc[7]) acc[7] = nf[0];
    }
    const bool is_max[NPART] = {false, false, false, false, false, false, false, true};
    block_reduce_store<NPART>(acc, is_max, partials + (long long)blockIdx.x * NPART);
}

// plaquette_charge / project_to_plane (KZM_bulk_ldg.cu:481-541): -1 unusable, 0, 1
__device__ int plaquette3(const double* a, const double* b, const double* c, const double* d, int plane,
                          double thr, double cutoff) {
    const double* nf[4] = {a, b, c, d};
    for (int t = 0; t < 4; ++t)
        if (!isfinite(nf[t][0])) return -1;
    for (int t = 0; t < 4; ++t)
        if (!(nf[t][0] > thr)) return -1;
    double p[4];
    for (int t = 0; t < 4; ++t) {
        double u = plane == 1 ? nf[t][2] : nf[t][1];
        double w = plane == 0 ? nf[t][2] : nf[t][3];
        const double nrm = sqrt(u * u + w * w);
        if (!(nrm > 1e-8)) return -1;
        p[t] = 2.0 * atan2(w / nrm, u / nrm);
    }
    const double dsum = wrap_loop(p[1] - p[0]) + wrap_loop(p[2] - p[1]) + wrap_loop(p[3] - p[2]) + wrap_loop(p[0] - p[3]);
    const double s = 0.5 * (dsum / (2.0 * 3.14159265358979323846));
    return fabs(s) > cutoff ? 1 : 0;
}

// partial slots: defects, plaquettes used, line length
__global__ __launch_bounds__(256) void obs_plaquette3_kernel(const double* __restrict__ nf, ObsParams P, double thr,
                                                             double cutoff, double* __restrict__ partials) {
    const long long n = (long long)P.Nx * P.Ny * P.Nz;
    double acc[NPLAQ] = {0, 0, 0};
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const int k = int(v % P.Nz), j = int((v / P.Nz) % P.Ny), i = int(v / ((long long)P.Ny * P.Nz));
        const int ip = (i + 1) % P.Nx, jp = (j + 1) % P.Ny, kp = (k + 1) % P.Nz;
        auto at = [&](int a, int b, int c) { return nf + 4 * (((long long)a * P.Ny + b) * P.Nz + c); };
        const double *n000 = at(i, j, k), *n100 = at(ip, j, k), *n110 = at(ip, jp, k), *n010 = at(i, jp, k);
        const double *n001 = at(i, j, kp), *n101 = at(ip, j, kp), *n011 = at(i, jp, kp);
        int r = plaquette3(n000, n100, n110, n010, 0, thr, cutoff);
        if (r >= 0) { acc[1] += 1.0; if (r) { acc[0] += 1.0; acc[2] += P.dz; } }
        r = plaquette3(n000, n010, n011, n001, 1, thr, cutoff);
        if (r >= 0) { acc[1] += 1.0; if (r) { acc[0] += 1.0; acc[2] += P.dx; } }
        r = plaquette3(n000, n100, n101, n001, 2, thr, cutoff);
        if (r >= 0) { acc[1] += 1.0; if (r) { acc[0] += 1.0; acc[2] += P.dy; } }
    }
    const bool is_max[NPLAQ] = {false, false, false};
    block_reduce_store<NPLAQ>(acc, is_max, partials + (long long)blockIdx.x * NPLAQ);
}

// ---------------------------------------------------------------------------
// Slice pass (droplet): nematic field on a z-slice, then plaquette winding
// (QSR.cu:1114-1176), xi edges (QSR.cu:1042-1106) and the checkerboard sums
// (QSR.cu:2846-2873).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void obs_slice_nematic_kernel(const float* __restrict__ q, long long cs, ObsParams P,
                                                                int z, double* __restrict__ nf) {
    const long long n = (long long)P.Nx * P.Ny;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long v = t * P.Nz + z;  // t = i*Ny + j
    double Q[5], o[4];
    load_q(q + P.off, cs, v, Q);
    nematic(Q, false, o);
    for (int c = 0; c < 4; ++c) nf[4 * t + c] = o[c];
}

// partial slots: defects, used, xi_sum, xi_edges, even_sum, even_n, odd_sum, odd_n, all_sum, all_n
__global__ __launch_bounds__(256) void obs_slice_stats_kernel(const double* __restrict__ nf_def,
                                                              const double* __restrict__ nf_xi, ObsParams P,
                                                              const RowMask* __restrict__ obs_rows, int z_def,
                                                              double S_thr, double cutoff, double xi_thr,
                                                              int want_checker, double* __restrict__ partials) {
    const long long n = (long long)P.Nx * P.Ny;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double acc[NSLICE] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (t < n) {
        const int i = int(t / P.Ny), j = int(t % P.Ny);
        auto ok = [](const double* f, double thr) {
            return isfinite(f[0]) && isfinite(f[1]) && isfinite(f[2]) && (f[0] > thr);
        };
        const double* a = nf_def + 4 * t;
        if (i < P.Nx - 1 && j < P.Ny - 1) {
            const double* b = nf_def + 4 * (t + P.Ny);
            const double* c = nf_def + 4 * (t + P.Ny + 1);
            const double* d = nf_def + 4 * (t + 1);
            if (ok(a, S_thr) && ok(b, S_thr) && ok(c, S_thr) && ok(d, S_thr)) {
                const double p00 = 2.0 * atan2(a[2], a[1]), p10 = 2.0 * atan2(b[2], b[1]);
                const double p11 = 2.0 * atan2(c[2], c[1]), p01 = 2.0 * atan2(d[2], d[1]);
                const double dsum = wrap_floor(p10 - p00) + wrap_floor(p11 - p10) + wrap_floor(p01 - p11) +
                                    wrap_floor(p00 - p01);
                const double s = 0.5 * (dsum * (1.0 / (2.0 * 3.141592653589793238462643383279502884)));
                acc[1] = 1.0;
                if (fabs(s) > cutoff) acc[0] = 1.0;
            }
        }
        const double* x0 = nf_xi + 4 * t;
        if (ok(x0, xi_thr)) {
            const double p0 = 2.0 * atan2(x0[2], x0[1]);
            if (i < P.Nx - 1) {
                const double* x1 = nf_xi + 4 * (t + P.Ny);
                if (ok(x1, xi_thr)) {
                    const double dd = wrap_floor(2.0 * atan2(x1[2], x1[1]) - p0);
                    acc[2] += dd * dd;
                    acc[3] += 1.0;
                }
            }
            if (j < P.Ny - 1) {
                const double* x1 = nf_xi + 4 * (t + 1);
                if (ok(x1, xi_thr)) {
                    const double dd = wrap_floor(2.0 * atan2(x1[2], x1[1]) - p0);
                    acc[2] += dd * dd;
                    acc[3] += 1.0;
                }
            }
        }
        if (want_checker) {
            const RowMask om = obs_rows[P.rows_off + (long long)i * P.Ny + j];
            const double S = a[0];
            if (in_iv(z_def, om.v[0], om.v[1]) && isfinite(S) && S > 0.01) {
                acc[8] = S;
                acc[9] = 1.0;
                if (((i + j) & 1) == 0) { acc[4] = S; acc[5] = 1.0; }
                else { acc[6] = S; acc[7] = 1.0; }
            }
        }
    }
    const bool is_max[NSLICE] = {false, false, false, false, false, false, false, false, false, false};
    block_reduce_store<NSLICE>(acc, is_max, partials + (long long)blockIdx.x * NSLICE);
}

// Fixed-order reduction of nblk x nslot partials into out[nslot]; max_mask
// bit s marks max slots.  One block.
__device__ __forceinline__ void finalize_block(const double* __restrict__ partials, long long nblk, int nslot,
                                               unsigned max_mask, double* __restrict__ out) {
    __shared__ double red[256];
    for (int s = 0; s < nslot; ++s) {
        const bool mx = (max_mask >> s) & 1u;
        double acc = 0.0;
        for (long long b = threadIdx.x; b < nblk; b += 256) {
            const double x = partials[b * nslot + s];
            acc = mx ? fmax(acc, x) : acc + x;
        }
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] = mx ? fmax(red[threadIdx.x], red[threadIdx.x + w])
                                                       : red[threadIdx.x] + red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[s] = red[0];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void obs_finalize_kernel(const double* __restrict__ partials, long long nblk,
                                                           int nslot, unsigned max_mask, double* __restrict__ out) {
    finalize_block(partials, nblk, nslot, max_mask, out);
}

// finalize_block for n boxes, one block each: box m's partials [prefix[m], prefix[m+1]) into
// out + m * nslot.
__global__ __launch_bounds__(256) void obs_finalize_multi_kernel(const double* __restrict__ partials,
                                                                 const long long* __restrict__ prefix, int nslot,
                                                                 unsigned max_mask, double* __restrict__ out) {
    const int m = blockIdx.x;
    finalize_block(partials + prefix[m] * nslot, prefix[m + 1] - prefix[m], nslot, max_mask,
                   out + (long long)m * nslot);
}

// Nematic field of a whole box, AoS double {S, nx, ny, nz} (computeNematicFieldKernel, QSR.cu:1017-1025)
__global__ __launch_bounds__(256) void nematic_aos_kernel(const float* __restrict__ q, long long cs, long long off,
                                                          long long n, double* __restrict__ out) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double Q[5], o[4];
    load_q(q + off, cs, v, Q);
    nematic(Q, false, o);
    for (int c = 0; c < 4; ++c) out[4 * v + c] = o[c];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace qsro {

static unsigned blocks_for(long long n) { return unsigned((n + 255) / 256); }
// grid of the volume passes (grid-stride loops): enough blocks to fill the chip, few partials
static unsigned volume_blocks(long long n) { return std::min(blocks_for(n), 4096u); }
long long obs_volume_blocks(long long n) { return volume_blocks(n); }

hipError_t launch_obs_droplet(hipStream_t st, const float* q, long long cs, const ObsParams& P,
                              const RowMask* step_rows, const RowMask* obs_rows, double* partials, long long* nblk) {
    const long long n = (long long)P.Nx * P.Ny * P.Nz;
    *nblk = volume_blocks(n);
    hipLaunchKernelGGL(obs_droplet_kernel, dim3(volume_blocks(n)), dim3(256), 0, st, q, cs, P, step_rows, obs_rows,
                       partials);
    return hipGetLastError();
}

hipError_t launch_obs_periodic(hipStream_t st, const float* q, long long cs, const ObsParams& P, double* nf,
                               double* partials, long long* nblk) {
    const long long n = (long long)P.Nx * P.Ny * P.Nz;
    *nblk = volume_blocks(n);
    hipLaunchKernelGGL(obs_periodic_kernel, dim3(volume_blocks(n)), dim3(256), 0, st, q, cs, P, nf, partials);
    return hipGetLastError();
}

hipError_t launch_obs_plaquette3(hipStream_t st, const double* nf, const ObsParams& P, double thr, double cutoff,
                                 double* partials, long long* nblk) {
    const long long n = (long long)P.Nx * P.Ny * P.Nz;
    *nblk = volume_blocks(n);
    hipLaunchKernelGGL(obs_plaquette3_kernel, dim3(volume_blocks(n)), dim3(256), 0, st, nf, P, thr, cutoff, partials);
    return hipGetLastError();
}

hipError_t launch_obs_slice_nematic(hipStream_t st, const float* q, long long cs, const ObsParams& P, int z, double* nf) {
    const long long n = (long long)P.Nx * P.Ny;
    hipLaunchKernelGGL(obs_slice_nematic_kernel, dim3(blocks_for(n)), dim3(256), 0, st, q, cs, P, z, nf);
    return hipGetLastError();
}

hipError_t launch_obs_slice_stats(hipStream_t st, const double* nf_def, const double* nf_xi, const ObsParams& P,
                                  const RowMask* obs_rows, int z_def, double S_thr, double cutoff, double xi_thr,
                                  int want_checker, double* partials, long long* nblk) {
    const long long n = (long long)P.Nx * P.Ny;
    *nblk = blocks_for(n);
    hipLaunchKernelGGL(obs_slice_stats_kernel, dim3(blocks_for(n)), dim3(256), 0, st, nf_def, nf_xi, P, obs_rows, z_def,
                       S_thr, cutoff, xi_thr, want_checker, partials);
    return hipGetLastError();
}

hipError_t launch_obs_droplet_multi(hipStream_t st, const float* q, long long cs, const ObsParams* Ps,
                                    const long long* prefix, int n, long long total_blocks, const RowMask* step_rows,
                                    const RowMask* obs_rows, double* partials, double* out) {
    if (n <= 0 || total_blocks <= 0) return hipSuccess;
    if (total_blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(obs_droplet_multi_kernel, dim3(unsigned(total_blocks)), dim3(256), 0, st, q, cs, Ps, prefix, n,
                       step_rows, obs_rows, partials);
    hipLaunchKernelGGL(obs_finalize_multi_kernel, dim3(unsigned(n)), dim3(256), 0, st, partials, prefix, NPART, 1u << 7,
                       out);
    return hipGetLastError();
}

hipError_t launch_obs_finalize(hipStream_t st, const double* partials, long long nblk, int nslot, unsigned max_mask,
                               double* out) {
    hipLaunchKernelGGL(obs_finalize_kernel, dim3(1), dim3(256), 0, st, partials, nblk, nslot, max_mask, out);
    return hipGetLastError();
}

hipError_t launch_nematic_aos(hipStream_t st, const float* q, long long cs, long long off, long long n, double* out) {
    hipLaunchKernelGGL(nematic_aos_kernel, dim3(blocks_for(n)), dim3(256), 0, st, q, cs, off, n, out);
    return hipGetLastError();
}

}  // namespace qsro
