/*
 * qsr_amd.h -- C ABI of the MI355X-native QSR Landau-de Gennes relaxation /
 * quench engine (libqsr_amd.so).
 *
 * The reference has no library API: its hot path is the chain of __global__
 * kernels declared at QSR.cuh:98-133 and launched per iteration from the host
 * loops in QSR.cu (sim_mode 1: QSR.cu:3220-3339, sim_mode 3: QSR.cu:2708-2964,
 * periodic bulk: KZM_bulk_ldg.cu:918-971).  This header replaces that chain with
 * a context object: the caller describes one box (or a batch of boxes), uploads
 * a field in the reference's interchange format (AoS double {Qxx,Qxy,Qxz,Qyy,Qyz},
 * index k + Nz*(j + Ny*i), QSR.cuh:22-44 / QSR.cu:286), advances it, and reads
 * observables and fields back.
 *
 * Conventions: every function returns 0 on success and a negative QSR_E* code
 * on failure (qsr_last_error() gives the message for the calling thread).  The
 * library never calls exit().  The context owns device memory; the caller owns
 * every host buffer.  One host thread per context; all work is ordered on the
 * context's HIP stream.  Plain pointers and sizes only.
 */
#ifndef QSR_AMD_H
#define QSR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QSR_OK 0
#define QSR_E_ARG (-1)       /* invalid argument / size mismatch */
#define QSR_E_HIP (-2)       /* HIP runtime failure (message names the call) */
#define QSR_E_GEOMETRY (-3)  /* droplet mask not representable by row intervals */
#define QSR_E_STATE (-4)     /* call out of order (e.g. step before upload) */
#define QSR_E_NODEVICE (-5)  /* no HIP device */

/* Temperature protocol of one box (QSR.cu:2437-2449, 2679-2693). */
#define QSR_PROTOCOL_CONSTANT 0 /* sim_mode=1 single temperature: T = T_high */
#define QSR_PROTOCOL_STEP 1     /* T_high for pre_equil_iters, then T_low */
#define QSR_PROTOCOL_RAMP 2     /* T_high, then linear to T_low over ramp_iters */

typedef struct qsr_params {
    /* grid (QSR.cu:1970-1975) */
    int32_t Nx, Ny, Nz;
    double dx, dy, dz;
    /* bulk LdG (QSR.cu:1978-1983); b > 0 is the reference's "-B/3 trQ^3" convention */
    double a, b, c, T_star;
    int32_t modechoice; /* 1 = std, 2 = ravnik (QSR.cu:1363-1378) */
    /* elastic (QSR.cu:2010, 2079-2082): kappa used only when L1 = L2 = L3 = 0 */
    double kappa, L1, L2, L3;
    /* weak anchoring W (J/m^2; 0 = strong, QSR.cu:2113) and viscosity gamma */
    double W, gamma;
    double dt;          /* time step (s); the caller applies the dt_max clamp */
    double q_norm_cap;  /* Frobenius cap (QSR.cu:2611); 0 = off */
    /* geometry: 0 = confined droplet (QSR.cu:144-162), 1 = periodic bulk box
     * without droplet (KZM_bulk_ldg.cu:351-469) */
    int32_t periodic;
    double radius_index;     /* droplet radius in voxels; <= 0 -> 0.5*(min N - 1) */
    double center_index[3];  /* droplet centre in voxel units; NaN -> (N-1)/2 */
    /* temperature protocol and shell order policy (QSR.cu:2233-2283) */
    int32_t protocol;        /* QSR_PROTOCOL_* */
    double T_high, T_low;    /* CONSTANT uses T_high */
    int32_t pre_equil_iters, ramp_iters;
    int32_t boundary_custom; /* 0 = equilibrium S_shell = max(S_eq(T), 0); 1 = custom */
    double boundary_S;       /* custom shell order */
    /* additive Gaussian noise per component per step, Philox4x32-10 keyed by
     * (seed, voxel, iteration); sigma = 0 disables (KZM_bulk_ldg.cu:441-454) */
    double noise_sigma;
    uint64_t seed;
} qsr_params;

/* Log-point observables of one box (the quench_log / free_energy_vs_iteration
 * columns, QSR.cu:2677, 2787-2882, 3218).  FP64, reduced on the GPU. */
typedef struct qsr_obs {
    int64_t iteration;      /* index of the last completed step (-1 before any) */
    double T, S_shell;      /* temperature / shell order of that step */
    double bulk, elastic, anchoring, total; /* J (droplet) or J/m^3 (periodic) */
    double radiality;       /* <|n . r_hat|> (QSR.cu:994-1015) */
    double avg_S, max_S;    /* over sd < -2*min_spacing (droplet) or all sites */
    double checker_rel;     /* NaN when not evaluated (QSR.cu:2846-2873) */
    uint64_t defects;       /* plaquettes with |s| > cutoff */
    uint64_t plaquettes;    /* plaquettes used */
    double defect_density;  /* defects / plaquettes (NaN if none) */
    double xi_grad_proxy;   /* QSR.cu:2372-2410 (droplet) / KZM_bulk_ldg.cu:709-711 */
    uint64_t xi_edges;
    double defect_line_density; /* periodic box only (KZM_bulk_ldg.cu:706-708) */
} qsr_obs;

typedef struct qsr_obs_opts {
    int32_t z_slice;          /* defects_z_slice; < 0 -> Nz/2 (QSR.cu:2646) */
    double S_threshold;       /* defects_S_threshold (0.1) */
    double charge_cutoff;     /* defects_charge_cutoff (0.25) */
    int32_t xi_z_slice;       /* < 0 -> z_slice */
    double xi_S_threshold;    /* < 0 -> S_threshold */
    int32_t want_defects, want_xi; /* 0 skips that pass (columns -> NaN / 0) */
    int32_t skip_order;       /* droplet boxes: 1 skips the nematic pass, so avg_S, max_S,
                                 radiality and checker_rel are NaN (energies only, e.g. the
                                 early-stop check); with want_defects = want_xi = 0 the
                                 mid-plane pass is skipped too */
} qsr_obs_opts;

typedef struct qsr_box_info {
    uint64_t voxels;          /* Nx*Ny*Nz */
    uint64_t shell_voxels;    /* |sd| <= shell_half_thickness (QSR.cu:2160-2172) */
    uint64_t active_voxels;   /* voxels that can change value in a step */
    double shell_volume;      /* m^3 */
    double surface_area;      /* Thomsen estimate, m^2 (QSR.cu:164-172) */
    double nominal_thickness; /* 2*shell_half_thickness */
    double effective_thickness; /* delta_eff = V_shell / A (QSR.cu:174-185) */
    double S_eq;              /* at T_high */
    int64_t iteration;        /* steps taken so far */
} qsr_box_info;

/* Context lifetime.  qsr_create_batch packs n boxes (each its own geometry,
 * physics, protocol and seed) into one device field advanced by one launch per
 * step -- the batched quench-rate ladder.  Replaces the per-run cudaMalloc x18
 * block (QSR.cu:2202-2219) and __constant__ geometry upload (QSR.cu:2154). */
int qsr_create(const qsr_params* p, int device, void** ctx);
int qsr_create_batch(const qsr_params* p, int n_boxes, int device, void** ctx);
int qsr_destroy(void* ctx);
/* No counterpart (the reference exits on errors, QSR.cu:13-24): the calling
 * thread's last error message, and the number of visible HIP devices. */
const char* qsr_last_error(void);
int qsr_device_count(int* n);

/* Field transfer in the interchange format (replaces the H2D/D2H cudaMemcpy of
 * QSR.cu:2516, 3174, 2974, 3468).  n_voxels must equal the box's Nx*Ny*Nz. */
int qsr_upload_q(void* ctx, int box, const double* aos5, size_t n_voxels);
int qsr_download_q(void* ctx, int box, double* aos5, size_t n_voxels);
/* Nematic field {S, nx, ny, nz} per voxel, FP64 (computeNematicFieldKernel,
 * QSR.cu:1017-1025) */
int qsr_download_nematic(void* ctx, int box, double* out4, size_t n_voxels);

/* Advance every box by nsteps explicit-Euler iterations (one fused launch per
 * step; replaces QSR.cu:3223-3238 / 2713-2753 / KZM_bulk_ldg.cu:958-970). */
int qsr_step(void* ctx, int nsteps);
/* Same, but captured once into a hipGraph of `chunk` steps and replayed. */
int qsr_step_graph(void* ctx, int nsteps, int chunk);
/* Semi-implicit IMEX iteration: explicit step with -L_stab*lap moved to the
 * left, then jacobi_iters Jacobi sweeps (QSR.cu:734-849, 2721-2749). */
int qsr_step_imex(void* ctx, int nsteps, double L_stab, int jacobi_iters);

/* FP64 log-point observables (replaces reduce_energy, radiality, defect and xi
 * passes with their full-array D2H copies, QSR.cu:2323-2410, 2787-2874). */
int qsr_observables(void* ctx, int box, const qsr_obs_opts* opts, qsr_obs* out);
/* The total free energy of each listed box (qsr_observables' `total` with
 * energies only, bit for bit), all reductions enqueued before one host
 * synchronisation: the energy sum of the early-stop test (reduce_energy,
 * QSR.cu:2323-2341; the test, QSR.cu:3322-3335) for a batched context. */
int qsr_energies(void* ctx, const int* boxes, int n, double* totals);
/* qsr_energies for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference.  The same kernels, so the same totals bit for
 * bit, but the reduction partials live in a fixed device array of the library
 * (about 550 KB per GPU, held until the process ends) instead of buffers the
 * context allocates on first use; nothing is allocated or freed.  Limits: at
 * most 4096 boxes per call, droplet boxes only (a periodic box returns
 * QSR_E_ARG; use qsr_energies).  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (several host threads, several contexts on
 * one device) race; serialise them.  totals is a host array of n doubles. */
int qsr_energies_noalloc(void* ctx, const int* boxes, int n, double* totals);
/* Shell diagnostics of a box, the block QSR.cu prints before a run
 * (QSR.cu:2179-2187; geometry QSR.cu:164-185), and its step count. */
int qsr_box_info_get(void* ctx, int box, qsr_box_info* out);
/* Iteration counter of a box, the loop's `iter` that T(iter) follows
 * (QSR.cu:2679-2687). */
int qsr_set_iteration(void* ctx, int box, int64_t iteration);

/* Instrumentation for measurement: average duration (ms) of the last `qsr_step`
 * call's per-step launches, measured with hipEvents on the context stream. */
int qsr_last_step_ms(void* ctx, double* ms_per_step);
/* Autotuner: times every stencil-kernel variant and planes-per-tile choice
 * for trial_steps launches on the resident field (the trial launches write the
 * spare buffer only, so the state and iteration count do not change), three
 * sweeps over all choices, and keeps the one whose fastest sweep is fastest
 * (best_ms).  Results are bit-identical across choices. */
int qsr_autotune(void* ctx, int trial_steps, int* best_variant, int* best_ch, double* best_ms);
/* Hold (active = 0) or resume one box of a batch: a held box takes no part in
 * later steps (its tiles leave the launch, its iteration count and field stay
 * as they are), so a quench-rate ladder can retire finished droplets early.
 * No reference counterpart (the reference runs one droplet per process). */
int qsr_set_box_active(void* ctx, int box, int active);
/* The same for the n listed boxes at once (one synchronisation and one rebuild of
 * the tile tables, where n calls of qsr_set_box_active make n). */
int qsr_set_boxes_active(void* ctx, const int* boxes, int n, int active);
/* Set the stencil variant (tile shape: 0 = 64 k x 4 j rows, 1 = 32 k x 8 j rows)
 * and planes per tile explicitly. */
int qsr_set_tuning(void* ctx, int variant, int planes_per_tile);
/* Two explicit steps per launch (temporal blocking).  With on = 1 (the default)
 * a context whose boxes use L1 / kappa elasticity only, without noise and with
 * no box-face voxel that can change, runs its steps in pairs through a kernel
 * that reads the field once and writes it once per pair (the intermediate
 * field stays on chip); results equal single steps bit for bit.  on = 0 runs
 * every step on its own.  planes_per_tile > 0 sets the pair kernel's planes per
 * tile (qsr_autotune also tunes it).  No reference counterpart: it replaces
 * two passes of the reference's per-iteration kernel chain (QSR.cu:3223-3238). */
int qsr_set_step_fusion(void* ctx, int on, int planes_per_tile);
/* Whether qsr_step pairs steps for this context, the pair kernel's planes per
 * tile and the voxels its tiles cover (0 when it does not pair). */
int qsr_step_fusion_info(void* ctx, int* active, int* planes_per_tile, uint64_t* tile_voxels);
/* Instrumentation: run nsteps (<= 512) with hipEvents recorded on the
 * context stream around every launch; returns the average duration (ms) of the
 * stencil launch and of the box-face launch.  A context that pairs its steps
 * runs nsteps / 2 two-step launches, and main_ms is the duration of one of them
 * (two steps). */
int qsr_profile_steps(void* ctx, int nsteps, double* main_ms, double* face_ms);
/* Work description of one step: bytes moved per step if each active voxel is
 * read and written once (FP32 SoA, 40 B/voxel), voxels covered by launched
 * tiles, and box voxels. */
int qsr_step_work(void* ctx, uint64_t* active_voxels, uint64_t* tile_voxels, uint64_t* box_voxels);

/* Host physics shared by the CLI and bindings (QSR.cu:1363-1604). */
double qsr_S_eq(double a, double b, double c, double T, double T_star, int modechoice);
/* out: min_dx, S_scale, L_eff_total, D_total, dt_diff_total, L_eff_explicit,
 *      dt_diff, bulk_rate, dt_bulk, dt_max (estimate_stable_dt_window) */
int qsr_dt_estimate(double dx, double dy, double dz, double gamma, double a, double b, double c,
                    double T_star, int modechoice, double kappa, double L1, double L2, double L3,
                    double S_ref_used, double S0, double T_a, double T_b, int use_semi_implicit,
                    double L_stab, int jacobi_iters, double out[10]);
/* Warm start / resume: read a Q-tensor text file in the format of
 * saveQTensorToFile (QSR.cu:1818-1832; "i j k" + 9 tensor entries per row) into
 * AoS double {Qxx,Qxy,Qxz,Qyy,Qyz}, reading it as QSRvis.py:2274-2304 does
 * ('#' comments, off-diagonals averaged, rows outside the grid ignored, voxels
 * without a row = 0).  No reference counterpart in QSR.cu, which only writes
 * these files.  rows_read (may be NULL) receives the number of data rows. */
int qsr_load_qtensor_file(const char* path, int Nx, int Ny, int Nz, double* aos5_out, size_t n_voxels,
                          uint64_t* rows_read);
/* Write AoS double Q in the saveQTensorToFile format (QSR.cu:1818-1832). */
int qsr_save_qtensor_file(const char* path, const double* aos5, int Nx, int Ny, int Nz, size_t n_voxels);
/* Seeded initial states in the reference draw order (QSR.cu:1841-1904):
 * mode 0 random director, 1 radial, 2 isotropic + N(0, noise_amplitude). */
int qsr_init_field(const qsr_params* p, int mode, double S0, double noise_amplitude,
                   int64_t random_seed, double* aos5_out, size_t n_voxels);

/* ---------------------------------------------------------------------------
 * Learned warm-start step.  No counterpart in the reference: QSR.cu only has
 * the plain relaxation loop (QSR.cu:3126-3476), which stays the fallback and
 * the arbiter of every state.  A three-layer 3x3x3 convolution network (5 -> 32
 * -> 32 -> 5 channels, GELU, 36,357 parameters) predicts dQ = Q(k+N) - Q(k)
 * from Q(k) (field input), or modulates the increment u(k) = E(Q(k)) - Q(k)
 * of one explicit Euler step E: dQ = N u (1 + y) per voxel and component, the
 * network reading x = u / s (increment input; s the RMS of u per component over
 * the voxels a step can change, box faces excluded) or x = Q(k) (gain input);
 * step-size input reads u / s and applies one gain per voxel, N u (1 + y_0).
 * A zero network then is the plain step extrapolated N times.  It runs on BF16
 * matrix cores with FP32 accumulation, and is
 * trained with its own backward and Adam kernels on pairs of fields the plain
 * solver produced.  Parameters are one flat FP32 array:
 * W1[27][5][32] b1[32] W2[27][32][32] b2[32] W3[27][32][5] b3[5], tap
 * t = (di+1)*9 + (dj+1)*3 + (dk+1) for the neighbour (i+di, j+dj, k+dk),
 * y[v,co] = b[co] + sum_{t,ci} W[t][ci][co] x[v+off(t), ci] (zero outside).
 * ------------------------------------------------------------------------- */
enum {
    QSR_SURR_CHECKED = 0, /* keep the step only if F and the residual do not rise */
    QSR_SURR_FORCE = 1,   /* keep the step */
    QSR_SURR_PROPOSE = 2  /* measure, then restore the previous field */
};
enum {
    QSR_SURR_INPUT_FIELD = 0,     /* x = Q(k), dQ = y (the default) */
    QSR_SURR_INPUT_INCREMENT = 1, /* x = u(k) / s, dQ = N u (1 + y) */
    QSR_SURR_INPUT_GAIN = 2,      /* x = Q(k), dQ = N u (1 + y) */
    QSR_SURR_INPUT_STEP = 3       /* x = u(k) / s, dQ = N u (1 + y_0): one gain per voxel */
};
typedef struct qsr_surr_result {
    double F_before, F_after;     /* total free energy (J) before / with the step */
    double res_before, res_after; /* |Q' - Q|_2 of one explicit Euler step over the voxels a step can change */
    int32_t accepted;             /* 1 if the context now holds the stepped field */
    double ms_network;            /* device time of the network pass */
    double scale;                 /* the scale of the step kept (or of the last one tried) */
    int32_t tried;                /* proposals evaluated (0: nothing to propose, zero increment) */
    int64_t net_voxels;           /* voxels the network pass computed (those of its tiles) */
} qsr_surr_result;

int qsr_surr_num_params(size_t* n);
int qsr_surr_create(int device, void** net);
int qsr_surr_destroy(void* net);
/* uniform(-1/sqrt(fan_in), +) weights and biases, last layer x0.1, from seed */
int qsr_surr_init_params(void* net, uint64_t seed);
/* set resets the Adam state */
int qsr_surr_set_params(void* net, const float* params, size_t n);
int qsr_surr_get_params(void* net, float* params, size_t n);
int qsr_surr_save(void* net, const char* path);
int qsr_surr_load(void* net, const char* path);
/* One pass over a host batch: x5, target5 (may be NULL = 0) and pred5 are
 * (nbatch, X, Y, Z, 5) FP32, mask (may be NULL = 1) is (nbatch, X, Y, Z).
 * loss = sum mask * |pred - target|^2 / (5 nbatch X Y Z); grad receives its
 * gradient (flat layout), pred = y.  loss / grad need Z % 8 == 0.  Outputs
 * may be NULL. */
int qsr_surr_eval(void* net, int nbatch, int X, int Y, int Z, const float* x5, const float* target5,
                  const float* mask, double* loss, float* pred5, float* grad);
/* The same with the modulated prediction of increment and gain input:
 * pred = m5 (1 + y), m5 (nbatch, X, Y, Z, 5) FP32 (the normalised increment u / s);
 * under step-size input pred = m5 (1 + y_0), one gain per voxel. */
int qsr_surr_eval_gain(void* net, int nbatch, int X, int Y, int Z, const float* x5, const float* m5,
                       const float* target5, const float* mask, double* loss, float* pred5, float* grad);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected) with a given gradient. */
int qsr_surr_adam_step(void* net, const float* grad, double lr);
/* Training data: snapshot the current field of a context's box (and which of
 * its voxels a step can change) into device memory; returns its index. */
int qsr_surr_dataset_add(void* net, void* ctx, int box, int* snapshot);
/* The same, also recording the increment of one explicit Euler step from the
 * field (needed by the input snapshot of a pair under increment input). */
int qsr_surr_dataset_add_increment(void* net, void* ctx, int box, int* snapshot);
/* What the network reads and, for increment input, the horizon N its output is
 * scaled by (N > 0).  Training and application use the setting of the call. */
int qsr_surr_set_input(void* net, int kind, double horizon);
/* Which tiles of a box the application's network pass computes (a tile: 12 rows
 * x 32 voxels of k x up to 16 planes): QSR_SURR_TILES_DROPLET (the default)
 * those within the network's reach, 2 voxels, of a voxel a step can change;
 * QSR_SURR_TILES_ALL every tile.  The proposals are the same either way. */
enum { QSR_SURR_TILES_DROPLET = 0, QSR_SURR_TILES_ALL = 1 };
int qsr_surr_set_tiles(void* net, int mode);
int qsr_surr_dataset_size(void* net, int* n);
int qsr_surr_dataset_clear(void* net);
/* iters Adam steps on batches of `batch` random crop^3 windows (crop % 8 == 0)
 * of the pairs (input snapshot, target snapshot) pairs[2p], pairs[2p+1]; the
 * target is their difference (field input), or the target is the difference
 * / (N s) and the prediction (u / s)(1 + y), u the input snapshot's recorded
 * increment, with the network reading u / s (increment and step-size input) or
 * the input snapshot's field (gain input).  A crop face inside its box leaves its 3-voxel
 * band unscored.  Each pair's voxels are weighted by 1 / max(ms, weight_floor),
 * ms the mean square of its target over the voxels a step can change, so the
 * small changes late in a relaxation count more (weight_floor >= 1 weights
 * every pair alike).  loss_per_iter (may be NULL) gets the mean over scored
 * voxels of weight * |pred - target|^2 / 5. */
int qsr_surr_train(void* net, const int* pairs, int npairs, int iters, int batch, int crop, double lr,
                   uint64_t seed, double weight_floor, double* loss_per_iter);
/* Propose Q + dQ for every voxel a step can change in a single-box context;
 * mode QSR_SURR_*.  The iteration counter is not advanced. */
int qsr_surr_apply(void* net, void* ctx, int mode, qsr_surr_result* out);
/* The same with the proposals Q + scales[i] dQ, i = 0 .. nscales-1 (scales > 0):
 * QSR_SURR_CHECKED keeps the first that passes the check (one network pass;
 * later proposals rescale the first), FORCE and PROPOSE use scales[0]. */
int qsr_surr_apply_ladder(void* net, void* ctx, int mode, const double* scales, int nscales, qsr_surr_result* out);
/* Batched contexts: the same for each listed box (uploaded, not held, listed
 * once), one network pass over all listed boxes (each layer one launch over
 * their tiles); each box is kept or restored on its own by
 * its own energy and residual, and unlisted boxes are untouched.  out[i] is
 * the result of boxes[i]; its ms_network is the device time of all listed
 * boxes' network passes and its net_voxels their voxel count. */
int qsr_surr_apply_boxes(void* net, void* ctx, int mode, const double* scales, int nscales, const int* boxes,
                         int nboxes, qsr_surr_result* out);

#ifdef __cplusplus
}
#endif
#endif /* QSR_AMD_H */
