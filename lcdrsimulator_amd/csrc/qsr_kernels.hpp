// qsr_kernels.hpp -- device-side data shared by the step and observable
// kernels and the context code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qsrk {

// Box flags
enum : int {
    F_PERIODIC = 1,   // wrap on every face, no droplet (every voxel evolves)
    F_ANCHOR = 2,     // Rapini-Papoular penalty on the shell (W != 0, QSR.cu:623-652)
    F_BC = 4,         // strong anchoring re-imposed on the shell (W <= 0, QSR.cu:851-867)
    F_SKIPSHELL = 8,  // shell voxels are not updated (W == 0, QSR.cu:705-706)
    F_NOISE = 16,     // additive Philox noise on evolving voxels
    F_CAP = 32,       // Frobenius-norm cap (QSR.cu:721-731)
    F_L2 = 64,
    F_L3 = 128,
    F_INACTIVE = 256,  // box held (qsr_set_box_active): no tiles, faces skipped
};

// Static description of one box, FP32 coefficients pre-multiplied by the
// mobility-step m = dt/gamma so the kernel computes dQ directly.
struct BoxDev {
    int Nx, Ny, Nz, flags;
    long long off;       // voxel offset of the box inside each packed SoA plane
    long long rows_off;  // row offset into the packed row-mask tables
    float cB, cC;        // m*B, m*C (A varies per step, see StepDyn)
    float lx, ly, lz;    // m*Lk/d^2, Lk = kappa when L1=L2=L3=0, else L1
    float cW;            // m*W/delta_eff
    float cap2;          // q_norm_cap^2
    float sigma;         // noise standard deviation per component per step
    float ci, cj, ck;    // droplet centre, index units
    float sx, sy, sz;    // surface-normal scales d/r^2 per axis (index units)
    // anisotropic elasticity (SI FP32, m-scaled products)
    float hx, hy, hz;    // 1/(2d)  central first difference
    float fx, fy, fz;    // 1/d     one-sided first difference (box faces)
    float ix2, iy2, iz2; // 1/d^2
    float i4xy, i4xz, i4yz; // 1/(4 d_a d_b)
    float mL2, mL3;      // m*L2, m*L3
    unsigned long long seed;
};

// Per step, per box: m*A(T), S_shell(T), iteration index for the noise counter.
struct StepDyn {
    float cA, S_shell;
    unsigned int iter, pad;
};

// Row-mask entries: 8 x int16 per row (i, j), see qsrh::RowMasks.
struct alignas(16) RowMask {
    short v[8];
};

// Tile shapes of the stencil and sweep kernels (256 threads, 4 waves per block),
// chosen per context by the autotuner ("variant"):
//   0: 64 k x 4 j rows, one wave per row (wider rows: fewer tiles along k)
//   1: 32 k x 8 j rows, a wave covers two rows of 32 (smaller in-plane halo, and
//      waves straddling a droplet surface carry fewer idle lanes)
constexpr int TILE_THREADS = 256;
constexpr int tile_k(int variant) { return variant == 1 ? 32 : 64; }
constexpr int tile_j(int variant) { return variant == 1 ? 8 : 4; }

// Two-step L1 kernel: 256 threads over a region of X2_TK k x X2_TJ j points, whose
// inner (X2_TK - 2) x (X2_TJ - 2) points are the tile's outputs.  Eight rows rather than
// 16: at 93 VGPRs five 256-thread blocks fit a CU (20 waves) where two 512-thread blocks
// did (16), and the 30 x 6 output tiles follow small droplets more closely.
constexpr int X2_TK = 32, X2_TJ = 8;
constexpr int X2_OUT_K = X2_TK - 2, X2_OUT_J = X2_TJ - 2;

}  // namespace qsrk

namespace qsro {

using qsrk::RowMask;

// FP64 parameters of one box for the log pass.
struct ObsParams {
    int Nx, Ny, Nz, periodic;
    long long off, rows_off;
    double dx, dy, dz;
    double cx, cy, cz, rx, ry, rz, min_spacing;
    double c2, c3, c4;         // bulk energy coefficients (QSR.cu:887-899) or A/2, B/3, C/4 (periodic)
    double kappa, L1, L2, L3;  // elastic
    int no_L;                  // L1 = L2 = L3 = 0 -> kappa
    double W_over_delta;       // 0 disables the anchoring energy
    double S_shell;
    int skip_order;            // 1: no nematic pass (<S>, max S, radiality)
};

constexpr int NPART = 8;     // volume partial slots
constexpr int NSLICE = 10;   // slice partial slots
constexpr int NPLAQ = 3;     // periodic plaquette slots

}  // namespace qsro


// Host launchers (defined in qsr_step.hip / qsr_obs.hip)
namespace qsrk {
// elastic: 0 = L1/kappa stencil; otherwise the anisotropic stencil with the terms
// of the mask compiled in (1 = L2, 2 = L3, 3 = both).  reverse != 0 runs the L1/kappa stencil's
// tile list back to front (callers alternate it with the buffer parity, so a launch first reads
// the tiles the previous one wrote last; the anisotropic stencil ignores it)
hipError_t launch_step(int elastic, int variant, bool noise, int ntiles, hipStream_t st, const float* qin, float* qout,
                       long long cs, const BoxDev* boxes, const int4* tiles, const StepDyn* dyn, int nboxes,
                       const RowMask* rows, int ch, const float4* imex, int reverse = 0);
// Two explicit L1 / kappa steps per launch (dyn1, dyn2: the two steps' per-box rows) over
// tiles of X2_OUT_K x X2_OUT_J outputs; only for contexts without noise, live box faces or
// 64-bit offsets (5 x cs floats below 4 GiB).  reverse != 0 runs the tile list back to front
// (callers alternate it with the buffer parity so each launch first reads what the last one wrote last)
hipError_t launch_step2(int ntiles, hipStream_t st, const float* qin, float* qout, long long cs, const BoxDev* boxes,
                        const int4* tiles, const StepDyn* dyn1, const StepDyn* dyn2, const RowMask* rows, int ch,
                        int reverse);
// face_list: (box, i, j, k) of every box-face voxel that can change (built on the host)
hipError_t launch_faces(long long nlive, hipStream_t st, const float* qin, float* qout, long long cs,
                        const BoxDev* boxes, const StepDyn* dyn, const int4* face_list, const RowMask* rows);
// One semi-implicit sweep over the step tiles.  flags: 2 x nsweeps words, zero
// before the first step; row alternates between steps.
hipError_t launch_jacobi(int variant, int ntiles, hipStream_t st, const float* rhs, const float* qold, float* qnew,
                         long long cs, const BoxDev* boxes, const int4* tiles, const RowMask* rows, int ch,
                         const float4* coef, unsigned* flags, int sweep, int nsweeps, int row);
hipError_t launch_face_copy(long long nlive, hipStream_t st, const float* src, float* dst, long long cs,
                            const BoxDev* boxes, const int4* face_list);
hipError_t launch_bc(long long nvox, hipStream_t st, float* q, long long cs, const BoxDev* boxes, const StepDyn* dyn,
                     const long long* vox_prefix, int nboxes, const RowMask* rows);
hipError_t launch_aos_to_soa(hipStream_t st, const double* aos, float* q0, float* q1, long long cs, long long off,
                             long long n);
hipError_t launch_soa_to_aos(hipStream_t st, const float* q, double* aos, long long cs, long long off, long long n);
}  // namespace qsrk

namespace qsro {
long long obs_volume_blocks(long long n_voxels);  // blocks (= partial records) of the volume passes
hipError_t launch_obs_droplet(hipStream_t st, const float* q, long long cs, const ObsParams& P,
                              const RowMask* step_rows, const RowMask* obs_rows, double* partials, long long* nblk);
hipError_t launch_obs_periodic(hipStream_t st, const float* q, long long cs, const ObsParams& P, double* nf,
                               double* partials, long long* nblk);
hipError_t launch_obs_plaquette3(hipStream_t st, const double* nf, const ObsParams& P, double thr, double cutoff,
                                 double* partials, long long* nblk);
hipError_t launch_obs_slice_nematic(hipStream_t st, const float* q, long long cs, const ObsParams& P, int z, double* nf);
hipError_t launch_obs_slice_stats(hipStream_t st, const double* nf_def, const double* nf_xi, const ObsParams& P,
                                  const RowMask* obs_rows, int z_def, double S_thr, double cutoff, double xi_thr,
                                  int want_checker, double* partials, long long* nblk);
hipError_t launch_obs_finalize(hipStream_t st, const double* partials, long long nblk, int nslot, unsigned max_mask,
                               double* out);
// The droplet volume pass and its finalize for n boxes in two launches: box m has parameters Ps[m]
// and owns partial records [prefix[m], prefix[m+1]) (prefix[m+1] - prefix[m] = obs_volume_blocks of
// its voxels, total_blocks = prefix[n]); its NPART sums go to out + m * NPART, bit for bit equal to
// launch_obs_droplet + launch_obs_finalize on the box alone.
hipError_t launch_obs_droplet_multi(hipStream_t st, const float* q, long long cs, const ObsParams* Ps,
                                    const long long* prefix, int n, long long total_blocks, const RowMask* step_rows,
                                    const RowMask* obs_rows, double* partials, double* out);
hipError_t launch_nematic_aos(hipStream_t st, const float* q, long long cs, long long off, long long n, double* out);
}  // namespace qsro
