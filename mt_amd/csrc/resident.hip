// resident.hip -- plain Langevin dynamics with each trajectory resident in
// one workgroup's LDS across many steps.
//
// One workgroup owns one trajectory.  LDS holds, per bead, the position and
// the three rotated offsets with the Euler angles packed into their spare
// lanes (64 B), plus the bead's neighbour lists compacted into a uint16 CSR
// (bonds: local index << 1 | flag; LJ: 16 x local index, the neighbour's
// byte offset among the positions) whose places are sized for the full lists, so a
// distance-filtered list and its later refill share one layout.  Each
// thread owns PT beads and keeps their two HybridTaus states in registers.
// A launch runs every step up to the next
// list-rebuild boundary (LJPairsUpdateFreq steps; 100 in the batched
// workload), after the rebuild kernel of lists.hip has refreshed the lists
// when one is due, so HBM sees the state once per launch instead of twice
// per step.  Per step:
//   forces + Euler-Maruyama for own beads, every operand from LDS
//   -> barrier -> own new state to LDS -> barrier.
// The arithmetic is bead_force / em_update, the code the streamed step runs.
// If a trajectory's lists do not fit the LDS left over, that workgroup reads
// them from the global k-major arrays instead (same results, slower).
#include "bead_physics.h"
#include "engine_internal.h"

namespace mte {

constexpr int RES_MAX_THREADS = 1024;   // one bead per thread
constexpr int RES_MULTI_THREADS = 768;  // several beads per thread: leaves room for more VGPRs
constexpr int RES_MAX_PT = 3;
constexpr int STATE_ARRAYS = 4;         // xyz|flags, n|fi, p1|theta, p2|psi
constexpr int HDR_INTS = 2;             // per bead: list offset, packed counts
constexpr int SCAN_INTS = RES_MAX_THREADS;
constexpr int FLAG_INTS = 2;            // per-step "a bead left its skin" flags, alternating by step parity
// Noise offload (two beads per thread): a second-pass bead owned by a wave
// of the first SIMD (wave % 4 == 0) draws its normals through the thread
// RNG_HELP places further on, which has no second-pass bead of its own; the
// helper draws them one step ahead into LDS (two buffers by step parity).
// The first SIMD carries one more second-pass wave than the others when a
// trajectory is not a multiple of 256 beads; this moves the HybridTaus and
// Box-Muller work of its second-pass beads to an idle SIMD.
constexpr int RNG_HELP = 6 * 64;
constexpr int RNG_SLOTS = 128;          // offloaded beads at most (two waves)
constexpr int RNG_F4 = 2 * RNG_SLOTS * 2;   // float4: [parity][slot][rx, ra]
__device__ __forceinline__ int rng_slot(int t) { return (t / 256) * 64 + (t % 64); }
// Whether thread t's second-pass bead draws through a helper (t + RNG_HELP).
__device__ __forceinline__ bool rng_offloaded(int t, int m2, int nthreads) {
    return (t / 64) % 4 == 0 && t < m2 && t + RNG_HELP < nthreads && t + RNG_HELP >= m2 && rng_slot(t) < RNG_SLOTS;
}
// LJ entries copied to LDS are those within LJ_CUT + 2 SKIN_HALF (+ margin)
// of the bead at the launch's first step; while every bead stays within
// SKIN_HALF of where it was then, no dropped pair can come below LJ_CUT.
constexpr float SKIN_HALF = 0.75f;
constexpr float FILTER_R = LJ_CUT + 2.f * SKIN_HALF + 0.01f;
// Minimum room for list entries per bead when sizing the plan.
constexpr int MIN_ENTRIES_PER_BEAD = 24;
// List entries a thread requests at once when it fills its bead's place.
constexpr int FILL_BATCH = 8;

// Packed counts: bit 0 harmonic present, bits 1-4 longitudinal, 5-9 lateral,
// 10-19 LJ.
__device__ __forceinline__ int pack_counts(int h, int nlon, int nlat, int nlj) {
    return (h ? 1 : 0) | (nlon << 1) | (nlat << 5) | (nlj << 10);
}

// Lists compacted in LDS: packed entries with local indices (LJ entries as
// byte offsets), the encoding bead_force reads through LdsAcc.
struct LdsLists {
    const uint16_t *e;
    int off, cnt;
    __device__ __forceinline__ int h() const { return cnt & 1; }
    __device__ __forceinline__ int harm() const { return h() ? (int)e[off] : -1; }
    __device__ __forceinline__ int nlon() const { return (cnt >> 1) & 15; }
    __device__ __forceinline__ int lon(int k) const { return e[off + h() + k]; }
    __device__ __forceinline__ int nlat() const { return (cnt >> 5) & 31; }
    __device__ __forceinline__ int lat(int k) const { return e[off + h() + nlon() + k]; }
    __device__ __forceinline__ int nlj() const { return (cnt >> 10) & 1023; }
    __device__ __forceinline__ int lj(int k) const { return e[off + h() + nlon() + nlat() + k]; }
};
// LJ entries in LDS hold the neighbour's byte offset in the position array
// (16 times its local index), so reading its position takes no scaling.
__device__ __forceinline__ float4 lj_xyz(const LdsAcc &A, const LdsLists &Ls, int k) {
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(A.x) + Ls.lj(k));
}

// Writes a bead's new state to LDS; returns (sin theta, cos theta, sin psi,
// cos psi) for the owner's next force evaluation.
__device__ __forceinline__ float4 store_state(float4 *lx, float4 *ln, float4 *l1, float4 *l2, int i, float4 x,
                                              float4 a, const LatOffsets &L) {
    const Rot R = make_rot(a.x, a.y, a.z);
    float3 v = R.apply(0.f, 0.f, 1.f);
    float4 n = make_float4(v.x, v.y, v.z, 0.f);
    v = R.apply(L.p1x, L.p1y, L.p1z);
    float4 p1 = make_float4(v.x, v.y, v.z, 0.f);
    v = R.apply(L.p2x, L.p2y, L.p2z);
    float4 p2 = make_float4(v.x, v.y, v.z, 0.f);
    n.w = a.x;   // fi
    p1.w = a.y;  // theta
    p2.w = a.z;  // psi
    lx[i] = x;
    ln[i] = n;
    l1[i] = p1;
    l2[i] = p2;
    return make_float4(R.st, R.ct, R.sp, R.cp);
}

// The step loop over steps [first, nsteps).  FILTERED: the LJ lists in LDS
// hold only the pairs within FILTER_R at step `first`; each mobile bead's
// displacement from its position then is checked every step, and once any
// bead of the trajectory has moved SKIN_HALF or more the loop returns the
// next step, for the caller to restore the full lists before that step's
// force evaluation (dropped pairs contribute exactly zero force until then,
// so the sums are those over the full lists) and go on unfiltered.
// Returns nsteps when every step ran.
template <int PT, bool FILTERED, class Lists, class MakeLists>
__device__ __forceinline__ int run_steps(const mte_params &P, const DevTopo &T, const LdsAcc &A, int t, float4 *lx,
                                         float4 *ln, float4 *l1, float4 *l2, const bool mob[PT], uint4 s0[PT],
                                         uint4 s1[PT], float4 trig[PT], const float *noise, int first, int nsteps,
                                         const LatOffsets &L, const MakeLists &make_lists, int *flags, float4 *rbuf,
                                         int own_slot, int help_slot, const int ib[PT]) {
    const int base = T.tbase[t], nb = T.nb;
    const float box = T.tbox[t];
    const float wall_z = P.is_wall ? T.zs[t] : 0.f;
    float3 xr[PT];
#pragma unroll
    for (int k = 0; k < PT; k++) xr[k] = FILTERED && mob[k] ? f3(lx[ib[k]]) : make_float3(0.f, 0.f, 0.f);
    for (int s = first; s < nsteps; s++) {
        float4 nx[PT], na[PT];
        bool moved = false;
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (!mob[k]) continue;
            const int i = ib[k];
            const int b = base + i;
            float4 x = lx[i];
            float4 a = make_float4(ln[i].w, l1[i].w, l2[i].w, 0.f);
            Own I = make_own_trig(x, a, ln[i], l1[i], l2[i], trig[k]);
            set_box(I, box);
            float F[6];
            const Lists Ls = make_lists(i, b);
            bead_force(P, A, Ls, i, wall_z, I, F);
            float4 rx, ra;
            if (noise) {
                const float *nz = noise + (size_t)s * 6 * nb + 6 * (size_t)b;
                rx = make_float4(nz[0], nz[1], nz[2], 0.f);
                ra = make_float4(nz[3], nz[4], nz[5], 0.f);
            } else if (k == 1 && own_slot >= 0) {   // drawn by the helper (PT == 2)
                rx = rbuf[((s & 1) * RNG_SLOTS + own_slot) * 2];
                ra = rbuf[((s & 1) * RNG_SLOTS + own_slot) * 2 + 1];
            } else {
                rx = normals4(s0[k]);
                ra = normals4(s1[k]);
            }
            em_update(P, F, rx, ra, x, a);
            nx[k] = x;   // w keeps the packed flags
            na[k] = a;
            if (FILTERED) {
                const float dx = x.x - xr[k].x, dy = x.y - xr[k].y, dz = x.z - xr[k].z;
                moved |= dx * dx + dy * dy + dz * dz >= SKIN_HALF * SKIN_HALF;
            }
        }
        if (PT == 2 && help_slot >= 0 && s + 1 < nsteps) {   // the owner's draws of the next step
            rbuf[(((s + 1) & 1) * RNG_SLOTS + help_slot) * 2] = normals4(s0[PT - 1]);
            rbuf[(((s + 1) & 1) * RNG_SLOTS + help_slot) * 2 + 1] = normals4(s1[PT - 1]);
        }
        if (FILTERED && moved) flags[s & 1] = 1;
        __syncthreads();
        const bool left = FILTERED && flags[s & 1] != 0;
        if (FILTERED && threadIdx.x == 0) flags[(s + 1) & 1] = 0;   // read at the next step, after its barrier
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (!mob[k]) continue;
            trig[k] = store_state(lx, ln, l1, l2, ib[k], nx[k], na[k], L);
        }
        __syncthreads();
        if (left) return s + 1;   // uniform: every thread read the same flag
    }
    return nsteps;
}

template <int PT, int MAXT>
__global__ void __launch_bounds__(MAXT)
    k_resident(mte_params P, DevTopo T, DevState S, uint4 *seeds, const float *noise, int nsteps, LatOffsets L,
               int stride, int entry_cap, const int *order) {
    extern __shared__ float4 lds[];
    // A rebuild that ran out of list capacity stops the dynamics: the state
    // and RNG streams stay at that rebuild's step, and the host reports the
    // error with the steps not taken (every thread returns before any
    // barrier).
    if (*(volatile const int *)T.overflow) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(T.frozen, nsteps);
        return;
    }
    float4 *lx = lds, *ln = lds + stride, *l1 = lds + 2 * stride, *l2 = lds + 3 * stride;
    float4 *rbuf = lds + STATE_ARRAYS * stride;
    int *hoff = reinterpret_cast<int *>(rbuf + RNG_F4);
    int *hcnt = hoff + stride;
    int *scan = hcnt + stride;
    int *flags = scan + SCAN_INTS;
    uint16_t *ent = reinterpret_cast<uint16_t *>(flags + FLAG_INTS);
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    const int nb = T.nb;
    if (threadIdx.x == 0) flags[0] = 0;
    stage_state(T, S, base, n, lx, ln, l1, l2);
    bool mob[PT];
    uint4 s0[PT], s1[PT];
    float4 trig[PT];
    int ib[PT];   // the bead of each of this thread's lane slots (order: immobile beads last)
    int lcnt[PT];   // its list counts, packed as in hcnt (requested with the bead's state)
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int slot = threadIdx.x + k * blockDim.x;
        const int i = slot < n ? order[base + slot] : slot;
        ib[k] = i;
        const int b = base + i;
        mob[k] = i < n && !T.fixed[b] && !T.extra[b];
        lcnt[k] = 0;
        if (mob[k])
            lcnt[k] = pack_counts(T.harm[b] >= 0, T.lon_n[b], T.lat_n[b], P.lj_on ? T.lj_n[b] : 0);
        trig[k] = make_float4(0.f, 1.f, 0.f, 1.f);
        if (mob[k]) {
            const float4 a = S.ang[b];
            sincos_fast(a.y, &trig[k].x, &trig[k].y);
            sincos_fast(a.z, &trig[k].z, &trig[k].w);
        }
        if (mob[k] && !noise) {
            s0[k] = seeds[b];
            s1[k] = seeds[b + nb];
        }
    }
    // noise offload roles (two beads per thread, drawn noise); the first
    // wave offloads too only when a later second-pass wave of the first SIMD
    // has work (its lanes past the mobile beads hold the immobile ones)
    const int m2 = n - (int)blockDim.x;
    // (flags[0] was cleared before the first barrier; the skin test clears it
    // again in fill before its first use)
    __syncthreads();
    if (PT == 2 && threadIdx.x >= 256 && mob[PT - 1]) flags[0] = 1;
    __syncthreads();
    const bool tail_busy = flags[0] != 0;
    auto offloaded = [&](int tt) {
        return PT == 2 && !noise && rng_offloaded(tt, m2, blockDim.x) && (tt >= 256 || tail_busy);
    };
    const int own_slot = offloaded(threadIdx.x) ? rng_slot(threadIdx.x) : -1;
    const int owner = (int)threadIdx.x - RNG_HELP;
    int help_slot = -1, help_b = -1;
    if (owner >= 0 && offloaded(owner)) {
        const int b = base + order[base + owner + blockDim.x];   // the owner's second-pass bead
        if (!T.fixed[b] && !T.extra[b]) {
            help_slot = rng_slot(owner);
            help_b = b;
            s0[PT - 1] = seeds[b];
            s1[PT - 1] = seeds[b + nb];
            rbuf[help_slot * 2] = normals4(s0[PT - 1]);   // step 0's draws (parity 0)
            rbuf[help_slot * 2 + 1] = normals4(s1[PT - 1]);
        }
    }
    __syncthreads();
    const LdsAcc A{lx, ln, l1, l2};
    const float box = T.tbox[t], ibox = inv_side(box);
    // LJ entries of bead i kept in LDS: all of them (rf2 == 0), or those
    // within sqrt(rf2) of it now.
    auto lj_keep = [lx, box, ibox](int i, int j, float rf2) {
        if (rf2 == 0.f) return true;
        const float3 d = min_image_pre(sub3(f3(lx[i]), f3(lx[j])), box, ibox);
        return d.x * d.x + d.y * d.y + d.z * d.z < rf2;
    };
    // Offsets of this trajectory's full lists in LDS, for the mobile beads
    // (the only ones whose forces are computed): the per-bead totals go
    // through a block scan into hoff.  Returns the number of entries.  Every
    // later fill writes at these offsets (a filtered list uses a prefix of
    // its bead's place), so the layout is computed once per launch.
    auto layout = [&]() {
        for (int i = threadIdx.x; i < n; i += blockDim.x) hoff[i] = 0;   // immobile beads use no lists
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (!mob[k]) continue;
            const int i = ib[k], hc = lcnt[k];
            hcnt[i] = hc;
            hoff[i] = (hc & 1) + ((hc >> 1) & 15) + ((hc >> 5) & 31) + (hc >> 10);
        }
        // Chunked exclusive scan: thread t owns beads [t*c, (t+1)*c); a wave
        // scans its lanes' chunk totals by shuffles, the first wave the
        // wave totals.
        const int c = (n + blockDim.x - 1) / blockDim.x;
        __syncthreads();
        int run = 0;
        const int lo = threadIdx.x * c, hi = min(n, lo + c);
        for (int i = lo; i < hi; i++) run += hoff[i];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
        int inc = run;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(inc, d, 64);
            if (lane >= d) inc += v;
        }
        if (lane == 63) scan[wv] = inc;
        __syncthreads();
        if (wv == 0) {
            int w = lane < nw ? scan[lane] : 0;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int v = __shfl_up(w, d, 64);
                if (lane >= d) w += v;
            }
            if (lane < nw) scan[lane] = w;
        }
        __syncthreads();
        const int total = scan[nw - 1];
        int acc = (wv ? scan[wv - 1] : 0) + inc - run;   // exclusive prefix of this chunk
        for (int i = lo; i < hi; i++) {
            const int v = hoff[i];
            hoff[i] = acc;
            acc += v;
        }
        __syncthreads();
        return total;
    };
    // Copy the lists into their places, of the LJ entries only those
    // lj_keep admits; hcnt's LJ field records the number copied.  A bead's
    // entries (harmonic, longitudinal, lateral, LJ: the full counts of the
    // layout) are requested FILL_BATCH at a time, each from its own array.
    // Each thread fills the places of its own mobile beads.
    int nent = 0;   // entries of the full lists (layout)
    auto fill = [&](float rf2) {
#pragma unroll
        for (int kk = 0; kk < PT; kk++) {
            if (!mob[kk]) continue;
            const int i = ib[kk], b = base + i;
            const int hc = hcnt[i];
            const int h = hc & 1, nlon = (hc >> 1) & 15, nlat = (hc >> 5) & 31;
            const int nbond = h + nlon + nlat;
            const int o0 = hoff[i];
            const int nall = (i + 1 < n ? hoff[i + 1] : nent) - o0;
            int o = o0, nlj = 0;
            for (int k0 = 0; k0 < nall; k0 += FILL_BATCH) {
                int v[FILL_BATCH];
#pragma unroll
                for (int u = 0; u < FILL_BATCH; u++) {
                    const int k = k0 + u;
                    const int *src = k < h          ? T.harm
                                     : k < h + nlon ? T.lon + (size_t)(k - h) * nb
                                     : k < nbond    ? T.lat + (size_t)(k - h - nlon) * nb
                                                    : T.lj + (size_t)(k - nbond) * nb;
                    v[u] = k < nall ? src[b] : 0;
                }
#pragma unroll
                for (int u = 0; u < FILL_BATCH; u++) {
                    const int k = k0 + u;
                    if (k < nbond) {
                        ent[o++] = (uint16_t)((((v[u] >> 1) - base) << 1) | (v[u] & 1));
                    } else if (k < nall && lj_keep(i, v[u] - base, rf2)) {
                        ent[o++] = (uint16_t)((v[u] - base) * (int)sizeof(float4));
                        nlj++;
                    }
                }
            }
            hcnt[i] = (hc & 1023) | (nlj << 10);
        }
        if (threadIdx.x == 0) flags[0] = flags[1] = 0;
        __syncthreads();
    };
    // The full lists decide between LDS and global lists; in LDS, the LJ
    // lists are filtered by distance when the list cutoff reaches beyond
    // FILTER_R (the full lists fit, so a refill always fits too).
    nent = layout();
    const bool use_lds = nent <= entry_cap;
    if (use_lds) {
        const bool filter = P.lj_on && P.ljpairscutoff > FILTER_R;
        fill(filter ? FILTER_R * FILTER_R : 0.f);
        auto mk = [&](int i, int) { return LdsLists{ent, hoff[i], hcnt[i]}; };
        int s = 0;
        if (filter) {
            s = run_steps<PT, true, LdsLists>(P, T, A, t, lx, ln, l1, l2, mob, s0, s1, trig, noise, 0, nsteps, L, mk,
                                              flags, rbuf, own_slot, help_slot, ib);
            if (s < nsteps) fill(0.f);   // a bead left its skin: the full lists from step s on
        }
        if (s < nsteps)
            run_steps<PT, false, LdsLists>(P, T, A, t, lx, ln, l1, l2, mob, s0, s1, trig, noise, s, nsteps, L, mk,
                                           flags, rbuf, own_slot, help_slot, ib);
    } else {
        auto mk = [&](int, int b) { return GlobalLists{T, b, base}; };
        run_steps<PT, false, GlobalLists>(P, T, A, t, lx, ln, l1, l2, mob, s0, s1, trig, noise, 0, nsteps, L, mk,
                                          flags, rbuf, own_slot, help_slot, ib);
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = base + i;
        float4 x = lx[i], fn = ln[i], f1 = l1[i], f2 = l2[i];
        x.w = 0.f;
        S.xyz[b] = x;
        S.ang[b] = make_float4(fn.w, f1.w, f2.w, 0.f);
        fn.w = f1.w = f2.w = 0.f;
        S.fn[b] = fn;
        S.fp1[b] = f1;
        S.fp2[b] = f2;
    }
#pragma unroll
    for (int k = 0; k < PT; k++) {
        if (mob[k] && !noise && !(k == 1 && own_slot >= 0)) {   // an offloaded bead's state is the helper's
            const int b = base + ib[k];
            seeds[b] = s0[k];
            seeds[b + nb] = s1[k];
        }
    }
    if (help_b >= 0) {
        seeds[help_b] = s0[PT - 1];
        seeds[help_b + nb] = s1[PT - 1];
    }
}

static size_t fixed_lds_bytes(int nmax) {
    return (size_t)STATE_ARRAYS * sizeof(float4) * nmax + (size_t)RNG_F4 * sizeof(float4) +
           (size_t)HDR_INTS * sizeof(int) * nmax + (size_t)(SCAN_INTS + FLAG_INTS) * sizeof(int);
}

// Launch plan with pt beads per thread (pt = 1 allows up to 1024 threads,
// 2 and 3 up to 768 so the multi-bead kernels keep their registers).
bool resident_plan_for(const mte_engine *e, int pt, ResidentPlan *plan) {
    *plan = ResidentPlan{};
    if (e->P.tea_on || pt < 1 || pt > RES_MAX_PT) return false;
    int nmax = 0;
    for (int t = 0; t < e->ntraj; t++) nmax = std::max<int>(nmax, (int)(e->off[t + 1] - e->off[t]));
    if (nmax > 4096) return false;   // uint16 list entries (LJ entries: 16 x the local index)
    int lds_limit = 0;
    if (hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) != hipSuccess)
        return false;
    const size_t fixed = fixed_lds_bytes(nmax);
    if (fixed + (size_t)MIN_ENTRIES_PER_BEAD * 2 * nmax > (size_t)lds_limit) return false;
    int threads = (nmax + pt - 1) / pt;
    threads = ((threads + 63) / 64) * 64;
    // One bead per thread also covers a trajectory longer than the largest
    // workgroup when every bead past it is immobile: the lane order puts the
    // mobile beads first, and an immobile bead needs its place in LDS only.
    if (pt == 1 && threads > RES_MAX_THREADS && e->res_mobile_max <= RES_MAX_THREADS) threads = RES_MAX_THREADS;
    // Two beads per thread run on 12 waves when the batch allows: the
    // second-pass beads then leave idle waves on the third SIMD for the
    // noise offload.
    if (pt == 2 && nmax > RES_MULTI_THREADS) threads = RES_MULTI_THREADS;
    if (threads > (pt == 1 ? RES_MAX_THREADS : RES_MULTI_THREADS)) return false;
    plan->threads = threads;
    plan->per_thread = pt;
    plan->max_beads = nmax;
    plan->lds_bytes = (size_t)lds_limit & ~(size_t)15;
    plan->entry_cap = (int)std::min<size_t>((plan->lds_bytes - fixed) / 2, 1u << 30);
    return true;
}

// Default plan: the fewest beads per thread the batch allows.
bool resident_fits(const mte_engine *e, ResidentPlan *plan) {
    for (int pt = 1; pt <= RES_MAX_PT; pt++)
        if (resident_plan_for(e, pt, plan)) return true;
    return false;
}

template <int PT, int MAXT>
static hipError_t launch_pt(mte_engine *e, int nsteps, const float *noise) {
    const ResidentPlan &pl = e->rplan;
    hipError_t s = hipFuncSetAttribute((const void *)k_resident<PT, MAXT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds_bytes);
    if (s != hipSuccess) return s;
    hipLaunchKernelGGL((k_resident<PT, MAXT>), dim3(e->ntraj), dim3(pl.threads), pl.lds_bytes, e->stream, e->P,
                       e->T, e->S[e->cur], e->seeds, noise, nsteps, e->L, pl.max_beads, pl.entry_cap, e->res_order);
    return hipGetLastError();
}

hipError_t launch_resident(mte_engine *e, int nsteps, const float *noise) {
    const ResidentPlan &pl = e->rplan;
    switch (pl.per_thread) {
    case 1: return launch_pt<1, RES_MAX_THREADS>(e, nsteps, noise);
    case 2: return launch_pt<2, RES_MULTI_THREADS>(e, nsteps, noise);
    case 3: return launch_pt<3, RES_MULTI_THREADS>(e, nsteps, noise);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace mte
