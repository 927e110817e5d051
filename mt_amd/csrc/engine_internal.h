// engine_internal.h -- the engine object and the kernel launchers that the
// C ABI (engine.hip) drives.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <string>
#include <vector>

#include "device_common.h"

namespace mte {

struct TimingSlot {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
    double total_ms = 0.0;
    int64_t launches = 0;
};

struct TeaBuffers {
    float4 *ci = nullptr;      // per bead: row sums of squares of D_hat (x, y, z)
    float *eps = nullptr;      // per bead: sum of all D_hat entries
    float *beta = nullptr;     // per trajectory
    double *stats = nullptr;   // per trajectory: eps, beta (FP64)
    int *status = nullptr;     // first capricious failure (1 + traj) or 0
    float4 *R = nullptr;       // per bead: scaled random force
    int4 *units = nullptr;     // pair work units: traj, i0, j0, chunk index
    int nunits = 0;
    int unit_chunk = 0;        // j-chunk length the units were built for (the pair kernel's CHUNK)
    int *nchunks = nullptr;    // per trajectory
    int max_chunks = 0;
    float4 *partial = nullptr; // max_chunks x nb pair partial sums
    // MTE_TEA_MFMA: 64-bead tiles of the spatial order (trajectory, first
    // position), and per tile the bf16 hi/lo moment tables [2][64 columns]
    // [64 beads], the centre and radius, and the bead records [4][64] and
    // indices [64] (k_tea_moments); the units are then (tile I, tile J,
    // slot I, slot J)
    int2 *tiles = nullptr;
    int ntiles = 0;
    void *mom = nullptr;
    float4 *ctr = nullptr;
    float4 *rec = nullptr;
    int *rbead = nullptr;
    // epsilon with each pair evaluated once (k_tea_eps_sym): work units
    // (traj, I group, J group >= I group) of 1024-bead groups, and the
    // partial sums [side][group][5][nb]; null when they would not fit
    int4 *eps_units = nullptr;
    int eps_nunits = 0;
    int eps_ngrp = 0;          // most groups of one trajectory
    float *eps_part = nullptr;
};

// Beads per group of the one-pass epsilon kernel (16 chunks of 64) and the
// largest partial-sum buffer it may take.
constexpr int EPS_GROUP = 1024;
constexpr size_t EPS_PART_MAX_BYTES = size_t(6) << 30;

// j-chunk length of the TEA pair kernel.
constexpr int TEA_CHUNK_BEADS = 128;

// Largest trajectory that k_rebuild_order sorts spatially (larger ones keep
// the index order); k_tea_pairs_mfma needs the spatial order.
constexpr int REBUILD_ORDER_MAX = 2048;

// Hashed cell lists (cells.hip): per trajectory a power-of-two bucket table.
struct CellTables {
    const int *toff = nullptr;   // trajectory -> first bucket
    const int *tsize = nullptr;  // trajectory -> bucket count (power of two)
    int *count = nullptr;        // per bucket
    int2 *range = nullptr;       // per bucket: first slot, bead count
    int4 *cell = nullptr;        // per bead: cell coordinates, bucket
    int *rank = nullptr;         // per bead: its place among the beads of its bucket
    int *sorted = nullptr;       // beads in bucket order
    float4 *spos = nullptr;      // per slot: position of the bead there, w = its index (bit 31: extra)
    int *bsum = nullptr;         // scan partials, one per CELL_SCAN_SPAN buckets
    int total = 0;               // buckets over all trajectories
};
// Buckets per block of the global bucket scan (cells.hip).
constexpr int CELL_SCAN_SPAN = 4096;

// MLP-HI network and work buffers (mlp.hip).
constexpr int MLP_NN = 32, MLP_IN = 96, MLP_H = 128, MLP_OUT = 9;
constexpr int MLP_SEARCH_BLOCK = 64;   // beads per workgroup of the neighbour search
constexpr int MLP_ERR_PARTS = 16;      // workgroups per trajectory of the judge's error sums
constexpr int MLP_NPARAM = MLP_IN * MLP_H + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_H * MLP_OUT + MLP_OUT;
// mte_mlp_forward_noalloc: rows per call (the fixed staging array's size)
// and the device ordinals whose array address is kept.
constexpr int MLP_NOALLOC_ROWS = MTE_MLP_NOALLOC_ROWS;
constexpr int MLP_NOALLOC_MAX_DEVICES = 64;
struct MlpBuffers {
    float *w = nullptr, *wi = nullptr;        // FP32 master weights, bf16-rounded copy for inference
    uint4 *img = nullptr;                     // inference image: bf16 MFMA fragments + biases (mlp_image_bytes())
    float *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    float *feat = nullptr, *targ = nullptr, *ystep = nullptr;   // per bead
    uint8_t *fblk = nullptr;   // per bead: 24-float blocks of `feat` the step's search wrote (the rest are zeros)
    int2 *blocks = nullptr;    // neighbour-search workgroups: (trajectory, 64-bead block), round robin over trajectories
    int nblocks = 0;
    double *err = nullptr;                    // per trajectory: squared error, squared label norm
    double *err_part = nullptr;               // their partial sums, MLP_ERR_PARTS per trajectory
    int *nbad = nullptr;                      // beads whose prediction failed the Cholesky check
    // training work space for up to `cap` samples
    int cap = 0;
    float *x = nullptr, *t = nullptr, *hid = nullptr, *y = nullptr, *g4 = nullptr, *ga = nullptr, *gb = nullptr;
    double *partial = nullptr;
    int adam_t = 0;
    bool ready = false;                       // weights uploaded or trained
};

// Resident-mode launch plan: trajectories grouped by workgroup.
struct ResidentPlan {
    int threads = 0;        // workgroup size
    int per_thread = 0;     // beads per thread
    int max_beads = 0;      // largest trajectory
    size_t lds_bytes = 0;
    int entry_cap = 0;      // uint16 list entries that fit beside the state
};

}  // namespace mte

struct mte_engine {
    mte_params P{};
    mte_options O{};
    int nb = 0, ntraj = 0;
    int nmax = 0;            // beads in the largest trajectory
    bool periodic = false;   // some trajectory has a periodic box
    std::vector<int64_t> off;
    int device = 0;
    int num_cus = 1;
    hipStream_t stream = nullptr;
    int mode = MTE_MODE_STREAMED;

    mte::DevTopo T{};
    mte::DevState S[2]{};
    int cur = 0;
    mte::LatOffsets L{};

    uint4 *seeds = nullptr;
    float *noise = nullptr;
    int64_t noise_steps = 0, noise_cursor = 0;
    float4 *fxyz = nullptr, *fang = nullptr;
    double *en_bead = nullptr, *en_traj = nullptr;
    int2 *tiles = nullptr;
    int ntiles = 0;
    int *rb_order = nullptr;   // per trajectory: local bead handled at each rebuild position (spatial order)
    int *res_order = nullptr;  // per trajectory: local bead at each resident lane slot (immobile beads last)
    std::vector<uint8_t> fixed_h;   // host copy of the fixed flags (the resident lane order)
    int res_mobile_max = 0;          // most mobile beads in one trajectory (the resident plan)
    bool any_extra = false;          // some bead is flagged extra (TEA's epsilon weights are needed)
    int *mtl_counts = nullptr, *mtl_flags = nullptr;   // mte_mt_length: per trajectory, per bead
    size_t forces_lds = 0;     // > 0: forces by k_forces_lds with this much dynamic LDS
    int rb_traj_threads = 0;   // > 0: list rebuilds by k_rebuild_traj with this workgroup size
    size_t rb_traj_lds = 0;    // and this much dynamic LDS
    // {overflow, frozen} copied by k_publish_flags at the end of a run into
    // pinned, device-mapped host memory, so that the run's one stream
    // synchronisation also delivers them (flags_dev is its device address).
    int *flags_host = nullptr, *flags_dev = nullptr;
    mte::TeaBuffers tea;
    mte::CellTables cells;
    bool use_cells = false;
    mte::MlpBuffers mlp;
    float *noalloc = nullptr;   // this device's staging array of mte_mlp_forward_noalloc
    int hi_mode = 0;            // MTE_HI_*
    int mlp_check_every = 100;
    float mlp_tol = 0.02f;
    bool mlp_fallback = false;
    std::vector<double> mlp_last_err;
    int64_t mlp_bad = 0;
    mte::ResidentPlan rplan;
    mte_tuning tune{};   // current launch shapes (mte_autotune / mte_set_tuning)

    int64_t step = 0;
    int64_t lists_step = -1;
    bool lists_valid = false;
    bool tea_valid = false;

    bool timing = false;
    // Captured step sequences of the streamed path (list rebuilds run before
    // them), keyed by (epsilon/MLP-check phase, length, buffer parity).
    std::map<std::tuple<int64_t, int64_t, int>, hipGraphExec_t> graphs;
    bool use_graphs = true;
    std::map<std::string, mte::TimingSlot> timers;
    std::vector<hipEvent_t> ev_pool;   // timing events whose pair was read, for reuse
    std::vector<void *> allocations;
    std::string err;
};

namespace mte {

// Launchers (kernels live in forces.hip, lists.hip, tea.hip, resident.hip).
hipError_t launch_frames(mte_engine *e, int buf);
hipError_t launch_step_plain(mte_engine *e, const float *noise_step);
hipError_t launch_forces(mte_engine *e);
hipError_t launch_forces_tea(mte_engine *e, const float *noise_step);
hipError_t launch_energies(mte_engine *e);
hipError_t launch_rebuild(mte_engine *e);
hipError_t launch_rebuild_order(mte_engine *e);
size_t forces_lds_bytes(int nmax);
hipError_t launch_mt_length(mte_engine *e, float rad_hi, float rad_lo, float cos_min, int apply);
bool rebuild_traj_plan(int nmax, int lds_limit, int *threads, size_t *lds);
hipError_t launch_rebuild_cells(mte_engine *e);
// The 12 nm cell tables of the MLP-HI feature search, then the search
// itself (which builds the tables first unless build_tables is false).
hipError_t launch_mlp_cells(mte_engine *e);
// fblk null: all 96 features of every bead are written; else only the
// 24-float blocks holding neighbours, their number in fblk.
hipError_t launch_mlp_local(mte_engine *e, float *feat, float *targ, bool build_tables = true,
                            uint8_t *fblk = nullptr);
hipError_t launch_mlp_forward(mte_engine *e, const float *feat, int n, float *out);
// Address of the device's fixed staging array for mte_mlp_forward_noalloc
// (hipGetSymbolAddress once per device; call with `device` current).
hipError_t noalloc_scratch_init(int device, float **ptr);
hipError_t launch_mlp_step(mte_engine *e, const float *noise_step, float *yout);
hipError_t launch_mlp_error(mte_engine *e, const float *y, const float *t);
hipError_t launch_mlp_grad(mte_engine *e, const float *feat, const float *targ, int n);
hipError_t launch_mlp_adam(mte_engine *e, float lr, int t);
hipError_t launch_mlp_round(mte_engine *e);
hipError_t launch_mlp_pack(mte_engine *e);
size_t mlp_image_bytes();
hipError_t build_cells(mte_engine *e, float side, int K);
float candidate_radius(const mte_params &P);
hipError_t launch_tea_eps(mte_engine *e);
hipError_t tea_build_eps_units(mte_engine *e);
hipError_t upload_resident_order(mte_engine *e, const uint8_t *extra);
// Most mobile (neither fixed nor extra) beads in one trajectory under these extra flags.
int resident_mobile_max(const mte_engine *e, const uint8_t *extra);
hipError_t launch_tea_pairs(mte_engine *e);
hipError_t launch_tea_finish(mte_engine *e, const float *noise_step);
hipError_t launch_resident(mte_engine *e, int nsteps, const float *noise);
bool resident_fits(const mte_engine *e, ResidentPlan *plan);
// Resident plan with `pt` beads per thread; false if the batch cannot use it.
bool resident_plan_for(const mte_engine *e, int pt, ResidentPlan *plan);
// TEA pair work units and partial buffers for j-chunks of `chunk` beads.
hipError_t tea_build_units(mte_engine *e, int chunk);

}  // namespace mte
