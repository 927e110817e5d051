/*
 * mt_engine.h -- C ABI of the MI355X (gfx950) engine for MADDY's per-step
 * Langevin hot path: monomer forces, pair-list rebuilds, the Euler-Maruyama
 * integrator with HybridTaus noise, TEA hydrodynamics and the learned MLP-HI
 * correction, for a batch of trajectories of uneven sizes (CSR offsets).
 *
 * What each entry point replaces in the reference (paths relative to the
 * reference repository root):
 *   mte_create / mte_destroy   initIntegration / deleteIntegration,
 *                              src/compute_cuda.cu:977-1123, plus
 *                              initTeaIntegrator / deleteTeaIntegrator,
 *                              src/bdhitea.cu:13-55, and initRand,
 *                              src/HybridTaus.cu:21-30
 *   mte_run                    the per-step loop of compute(),
 *                              src/compute_cuda.cu:1137-1240 (LJ_kernel,
 *                              pairs_kernel, compute_kernel, updateTea +
 *                              integrateTea or integrate_kernel), between two
 *                              host events
 *   mte_set_state/get_state    the cudaMemcpy of d_r, :1013-1014, :1177, :1204
 *   mte_set_flags              the updater's re-uploads of extra / gtp /
 *                              on_tubule_cur, :1157, :1190, :1202
 *   mte_compute_energies       energy_kernel + host sums, :676-911 and
 *                              src/updater.cpp:3-43
 *   mte_compute_forces         compute_kernel, :32-525 (forces only)
 *   mte_get_lists              the device lists pairs_kernel / LJ_kernel
 *                              fill, :527-674, :913-940
 *   mte_tea_stats              tea.h_epsilon / h_beta_ij, src/bdhitea.cu:57-118
 *   mte_tea_ci                 tea.d_ci, src/bdhitea_kernel.cu:61-101
 *   mte_get_rng / mte_set_rng  ht.d_seeds, src/HybridTaus.cu:18-30
 *
 * Conventions: every call returns MTE_OK (0) or an MTE_ERR_* code and never
 * exits the process; mte_last_error() gives the message.  The caller owns all
 * host buffers; the engine owns device memory.  Calls on one engine are not
 * thread-safe.  mte_run is asynchronous inside and synchronous at return.
 */
#ifndef MT_ENGINE_H
#define MT_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MTE_OK 0
#define MTE_ERR_ARG 1       /* bad argument or topology */
#define MTE_ERR_HIP 2       /* HIP runtime error */
#define MTE_ERR_OVERFLOW 3  /* a neighbour list exceeded its capacity */
#define MTE_ERR_TEA 4       /* TEA capricious check (src/bdhitea.cu:89-107) */
#define MTE_ERR_STATE 5     /* call not valid in the current state */
#define MTE_ERR_NOISE 6     /* injected noise exhausted */

/* Same layout as the reference's Coord, src/mt.h:63-71. */
typedef struct { float x, y, z, fi, theta, psi, w; } mte_coord;

/* Same layout as the reference's Energies, src/mt.h:92-100. */
typedef struct { double U_harm, U_long, U_lat, U_psi, U_fi, U_teta, U_lj; } mte_energies;

/* Derived constants the hot path reads (subset of Parameters,
 * src/parameters.h:246-318; derivation src/preparator.cpp:124-221). */
typedef struct {
    float dt, temp, gammaR, gammaTheta, varR, varTheta, alpha, freeze_temp;
    float C, B_psi, B_fi, B_theta, psi_0, fi_0, theta0_gtp, theta0_gdp;
    float A_long, D_long, A_lat, D_lat, seam_coeff;
    int32_t barrier;
    float a_barr_long, r_barr_long, w_barr_long, a_barr_lat, r_barr_lat, w_barr_lat;
    int32_t lj_on;
    float ljscale, ljsigma6, ljpairscutoff;
    int32_t ljpairsupdatefreq;
    int32_t is_wall;
    float rep_h, rep_r, rep_eps, rep_leftborder;
    int32_t is_assembly;
    int32_t tea_on;
    float tea_a;
    int32_t tea_epsilon_freq, tea_capricious;
    float tea_epsmax;
    /* The reference's compile-time REGULARIZATION (src/compute_cuda.cu:482-486)
     * as a run-time switch: each LJ pair force is capped at gammaR r_mon / dt
     * (r_mon = 2 nm, src/mt.h:24), the force that moves a bead by r_mon in one
     * step.  0 = off, as the reference builds by default. */
    int32_t lj_regularize;
} mte_params;

/* Topology of a CSR batch.  Trajectory t owns global beads
 * [traj_offsets[t], traj_offsets[t+1]).  List entries are LOCAL indices in
 * the reference's sign encodings (src/preparator.cpp:339-355 for harmonic,
 * with -1 for "no partner"; src/compute_cuda.cu:587-592 longitudinal;
 * :646-658 lateral, 999999 standing for j = 0).  The reference's
 * per-monomer arrays (harmonic, fixed, mon_type) are given per bead. */
typedef struct {
    int32_t ntraj;
    const int64_t *traj_offsets;        /* ntraj + 1 */
    const int32_t *harmonic_count;      /* B, 0 or 1 */
    const int32_t *harmonic;            /* B */
    const int32_t *longitudinal_count;  /* B: initial lists (used as-is when is_assembly == 0) */
    const int32_t *longitudinal;        /* B * 8 */
    const int32_t *lateral_count;       /* B */
    const int32_t *lateral;             /* B * 16 */
    const uint8_t *fixed, *extra;       /* B */
    const int32_t *mon_type, *gtp, *on_tubule; /* B */
    const float *wall_z;                /* ntraj: Parameters.zs */
    const float *box;                   /* ntraj: periodic cubic box side, 0 = open (the
                                           reference has open boundaries only); may be NULL */
} mte_topology;

#define MTE_MODE_AUTO 0
#define MTE_MODE_STREAMED 1  /* one thread per bead, state in HBM, per-step graph */
#define MTE_MODE_RESIDENT 2  /* one workgroup per trajectory, state in LDS across steps */

typedef struct {
    int32_t max_lj;   /* LJ list capacity per bead (the reference uses 256) */
    int32_t mode;     /* MTE_MODE_* */
    int32_t rseed;    /* HybridTaus seeding via ran2 (src/HybridTaus.cu:32-48) */
    int32_t device;   /* HIP device ordinal */
    int32_t list_method; /* MTE_LISTS_*: how streamed-mode rebuilds find candidates */
} mte_options;

#define MTE_LISTS_AUTO 0      /* cell lists for trajectories above 2048 beads */
#define MTE_LISTS_ALLPAIRS 1  /* every bead of the trajectory, LDS-tiled */
#define MTE_LISTS_CELLS 2     /* hashed cell lists, sort and scan */

typedef struct mte_engine mte_engine;

int mte_create(const mte_params *params, const mte_topology *topo, const mte_options *opt, mte_engine **out);
int mte_destroy(mte_engine *e);
/* Message of the last failing call on e (or of the last failed mte_create
 * when e is NULL). */
const char *mte_last_error(const mte_engine *e);

int mte_num_beads(const mte_engine *e, int64_t *nbeads, int32_t *ntraj);
int mte_set_state(mte_engine *e, const mte_coord *r);
int mte_get_state(mte_engine *e, mte_coord *r);
/* Any pointer may be NULL (left unchanged).  New extra flags reorder the
 * resident kernel's lanes (immobile beads last); a one-bead-per-thread plan
 * that no longer covers a trajectory's mobile beads becomes the default plan
 * (resident_per_thread in mte_get_tuning shows it). */
int mte_set_flags(mte_engine *e, const uint8_t *extra, const int32_t *gtp, const int32_t *on_tubule);

/* The updater's tubule classification (mt_length, src/updater.cpp:154-217)
 * on the engine's current state, without downloading it: a bead is on the
 * tubule when rad_lo < sqrt(x^2 + y^2) < rad_hi and cos(theta) > cos_min
 * (the reference: R_MT + R_THRES, 1, cos(ANG_THRES), src/mt.h:23,29-30).
 * counts[ntraj] receives each trajectory's count and on_tubule[B] (may be
 * NULL) the per-bead flags; with apply != 0 the flags also replace the
 * engine's on_tubule, as mte_set_flags(NULL, NULL, on_tubule) would. */
int mte_mt_length(mte_engine *e, float rad_hi, float rad_lo, float cos_min, int32_t apply, int32_t *counts,
                  int32_t *on_tubule);
int mte_set_step(mte_engine *e, int64_t step);
int mte_get_step(const mte_engine *e, int64_t *step);

/* Advance nsteps steps (global step counter continues).  Lists are rebuilt
 * at steps divisible by LJPairsUpdateFreq, TEA's epsilon/beta at steps
 * divisible by tea_epsilon_freq, as in the reference loop. */
int mte_run(mte_engine *e, int64_t nsteps);

/* Forces (and torques) at the current state, reference compute_kernel
 * semantics; f: B records, w = 0.  Rebuilds the lists first if the current
 * step is a rebuild step that has not run yet. */
int mte_compute_forces(mte_engine *e, mte_coord *f);
/* Per-bead energies (halved, energy_kernel convention) and/or per-trajectory
 * FP64 sums (ntraj x 7, Energies order).  Either pointer may be NULL. */
int mte_compute_energies(mte_engine *e, mte_energies *per_bead, double *per_traj);
int mte_rebuild_lists(mte_engine *e);
/* Lists in the reference encodings (local indices): lon B*8, lat B*16,
 * lj B*max_lj.  Any pointer may be NULL. */
int mte_get_lists(mte_engine *e, int32_t *lon_count, int32_t *lon, int32_t *lat_count, int32_t *lat,
                  int32_t *lj_count, int32_t *lj);

/* Test hook: nsteps x B x 6 floats of standard normals used instead of
 * HybridTaus (xyz = rforce(p).xyz, angles = rforce(p + B).xyz).  NULL
 * returns to HybridTaus. */
int mte_set_noise(mte_engine *e, const float *noise, int64_t nsteps);
/* HybridTaus states, 2 * B uint4 (xyz stream at p, angular at p + B). */
int mte_get_rng(mte_engine *e, uint32_t *seeds);
int mte_set_rng(mte_engine *e, const uint32_t *seeds);

/* Per-trajectory epsilon and beta of the last TEA refresh (FP64). */
int mte_tea_stats(mte_engine *e, double *eps, double *beta);
/* Per-bead row sums of squares of the normalised RPY tensor over j != i
 * (x, y, z rows; 3 floats per bead) from the last TEA refresh: the
 * reference's tea.d_ci (src/bdhitea_kernel.cu:61-101), from which
 * C_i = 1 / sqrt(1 + beta^2 ci) (eq. 19 of Geyer & Winter). */
int mte_tea_ci(mte_engine *e, float *ci);

/* Kernel timing with HIP events on the engine's stream. */
int mte_set_timing(mte_engine *e, int32_t on);
/* name: "step" (dominant kernel of plain mte_run), "rebuild", "forces",
 * "tea_pairs" (the TEA pair-sum kernel), "mlp_local" and "mlp_step" (the
 * MLP-HI feature search and the fused inference + integration kernel).
 * Returns summed milliseconds and launch count since the last reset. */
int mte_kernel_time(mte_engine *e, const char *name, double *total_ms, int64_t *launches);
int mte_reset_timing(mte_engine *e);
/* Which mode create() chose (MTE_MODE_STREAMED or MTE_MODE_RESIDENT). */
int mte_get_mode(const mte_engine *e, int32_t *mode);

/* ---------------------------------------------------------------- autotuning
 * Launch shapes of the engine's kernels (no counterpart in the reference,
 * whose launches are fixed at BLOCK_SIZE threads, src/Cuda.h:122-135).
 * A field set to 0 keeps the current choice. */
/* tea_chunk value: 64-bead chunks with each (i, j) RPY block evaluated once
 * for both beads (the other chunk values evaluate it for i and for j). */
#define MTE_TEA_SYMMETRIC 1
/* tea_chunk value: the pair sums on the matrix cores (open boundaries and
 * trajectories of at most 2048 beads; MTE_ERR_ARG otherwise): two scalar RPY
 * kernels per pair times per-tile moment tables in split bf16 products with
 * FP32 accumulation, each pair once for both beads, pairs closer than 6 nm
 * summed exactly on the VALU. */
#define MTE_TEA_MFMA 2

/* tea_eps values: each pair's epsilon terms evaluated once for both beads,
 * through a partial-sum buffer of 40 x ceil(n_max / 1024) bytes per bead
 * (the default when it stays within 6 GiB), or evaluated for each bead of
 * the pair (no buffer). */
#define MTE_TEA_EPS_ONEPASS 1
#define MTE_TEA_EPS_TWOSIDED 2

/* forces values (the force pass of TEA and MLP-HI steps): each workgroup
 * stages its bead's whole trajectory in LDS and reads neighbours there (only
 * for trajectories that fit), or neighbours are gathered from global memory
 * by workgroups of stream_block beads, each XCD taking a contiguous eighth of
 * the beads so that their trajectories' state stays in that XCD's L2. */
#define MTE_FORCES_LDS 1
#define MTE_FORCES_GLOBAL 2

typedef struct {
    int32_t resident_per_thread;  /* beads per thread of the resident kernel: 1, 2 or 3 */
    int32_t stream_block;         /* workgroup size of the streamed bead kernels: 64, 128 or 256 */
    int32_t tea_chunk;            /* j-chunk of the TEA pair kernel: 64, 128 or 256 beads, or
                                     MTE_TEA_SYMMETRIC or MTE_TEA_MFMA (0 without TEA) */
    int32_t cell_stencil;         /* cell-list rebuilds: 1 = 27 cells of side r, 2 = 125 cells of side r/2 */
    int32_t mode;                 /* MTE_MODE_RESIDENT or MTE_MODE_STREAMED (plain dynamics) */
    int32_t tea_eps;              /* TEA epsilon: MTE_TEA_EPS_ONEPASS or MTE_TEA_EPS_TWOSIDED (0 without TEA) */
    int32_t forces;               /* TEA / MLP-HI force pass: MTE_FORCES_LDS or MTE_FORCES_GLOBAL */
    int32_t reserved;
} mte_tuning;

int mte_get_tuning(const mte_engine *e, mte_tuning *t);
/* Applies the non-zero fields; MTE_ERR_ARG for a shape the batch cannot use. */
int mte_set_tuning(mte_engine *e, const mte_tuning *t);
/* Times every candidate shape that applies to the batch (trial_steps steps
 * each: for plain dynamics every resident plan against every streamed
 * workgroup size, which also picks the mode; then TEA chunks and cell
 * stencils one family after the other) and keeps the fastest.  Coordinates, RNG streams and the step counter are restored after
 * every trial, and lists and TEA's epsilon are recomputed at the next step;
 * call it where both are due anyway (e.g. right after mte_create, step 0) so
 * the tuned run is the untuned one.  MTE_ERR_STATE elsewhere, or with
 * injected noise pending.  `report` (may be NULL) receives one line per trial. */
int mte_autotune(mte_engine *e, int32_t trial_steps, mte_tuning *chosen, char *report, int64_t report_len);

/* ------------------------------------------------------------------ MLP-HI
 * The learned hydrodynamic correction has no counterpart in the reference.
 * Per bead it predicts the local HI tensor G_i = sum_j phi(r_ij) D_hat_ij
 * (RPY blocks of src/bdhitea_kernel.cu:38-58 over neighbours within 12 nm,
 * cosine taper from 8 nm) from the relative positions of the 32 nearest
 * neighbours with a 96-128-128-128-9 tanh MLP (bf16 weights, FP32
 * accumulation, on the matrix cores), and replaces TEA's per-step pair sums:
 *   dr = (dt/gammaR) (I + G) F + varR L xi,   L L^T = C (I + 2 beta G) C
 * with TEA's beta and per-bead C_i (src/bdhitea_kernel.cu:104-213),
 * refreshed every tea_epsilon_freq steps.  Every check_every steps the exact
 * G_i is computed for all beads; a per-trajectory relative error above the
 * tolerance hands the run back to TEA.
 * Weights: MTE_MLP_NPARAM floats, W1 (96x128, input-major), b1, W2, b2, W3,
 * b3, W4 (128x9), b4.  Requires tea_on parameters (tea_a, gammaR). */
#define MTE_MLP_NPARAM 46601
#define MTE_MLP_FEATURES 96
#define MTE_MLP_OUTPUTS 9
#define MTE_HI_DEFAULT 0  /* TEA when params.tea_on, else free draining */
#define MTE_HI_MLP 1      /* MLP-HI, falling back to TEA above the error tolerance */

int mte_mlp_set_weights(mte_engine *e, const float *w, int64_t n);
int mte_mlp_get_weights(mte_engine *e, float *w, int64_t n);
/* Features (B x 96) and exact labels A - I (B x 9) of the current state;
 * either pointer may be NULL. */
int mte_mlp_local(mte_engine *e, float *features, float *labels);
/* Inference with the bf16-rounded weights on n feature rows. */
int mte_mlp_forward(mte_engine *e, const float *features, int64_t n, float *out);
/* The same inference for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  No counterpart in the
 * reference.  mte_mlp_forward sizes its work space to the call and may
 * allocate it; this variant allocates nothing: features and outputs are
 * staged in a fixed device array of MTE_MLP_NOALLOC_ROWS rows (3.4 MB per
 * GPU, held until the process ends), and n above that limit is MTE_ERR_ARG.
 * The results equal mte_mlp_forward's bit for bit.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (two threads, or
 * two engines on one device) race. */
#define MTE_MLP_NOALLOC_ROWS 8192
int mte_mlp_forward_noalloc(mte_engine *e, const float *features, int64_t n, float *out);
/* Mean squared error and its gradient (MTE_MLP_NPARAM floats) with the FP32
 * weights over n samples. */
int mte_mlp_loss_grad(mte_engine *e, const float *features, const float *labels, int64_t n, double *loss,
                      float *grad);
/* Adam training on the device: `iters` steps over consecutive mini-batches
 * of `batch` samples (cycling through the n given); loss per step into
 * losses (may be NULL). */
int mte_mlp_train(mte_engine *e, const float *features, const float *labels, int64_t n, int32_t iters,
                  int32_t batch, float lr, double *losses);
/* Select the hydrodynamics of mte_run: MTE_HI_DEFAULT or MTE_HI_MLP, with the
 * exact-label check every check_every steps and the relative error above
 * which the engine falls back to TEA for the rest of the run. */
int mte_set_hi_mode(mte_engine *e, int32_t mode, int32_t check_every, float tolerance);
/* Relative error of the last check per trajectory, whether the fallback has
 * fired, and how many bead predictions failed the positivity check. */
int mte_mlp_status(mte_engine *e, double *last_error, int32_t *fallback, int64_t *bad_beads);

#ifdef __cplusplus
}
#endif
#endif
