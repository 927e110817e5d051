// forces.hip -- streamed-mode kernels: one thread per bead, bead state in
// HBM (float4 SoA, double-buffered), neighbour lists k-major so that a
// wavefront's list loads coalesce.
#include "bead_physics.h"
#include "engine_internal.h"

namespace mte {

// Workgroup size of k_frames; the per-step bead kernels take theirs from the
// engine's tuning (64, 128 or 256).
constexpr int STREAM_BLOCK = 256;

// Bead of this thread in a grid of 8 x nb8 workgroups (grid_xcd).  The
// dispatcher hands workgroup k to XCD k mod 8, so workgroup k takes the
// (k / 8)-th block of the (k mod 8)-th eighth of the beads: each XCD works
// on a contiguous eighth, whose trajectories' state then stays in that
// XCD's L2 while their neighbours are gathered.  The bead a thread computes
// does not change any result.
__device__ __forceinline__ int xcd_bead(int bs) {
    const int nb8 = (int)gridDim.x >> 3;
    return (((int)blockIdx.x & 7) * nb8 + ((int)blockIdx.x >> 3)) * bs + (int)threadIdx.x;
}

constexpr int FORCE_TILE = 256;   // k_forces_lds: beads per workgroup (the 256-bead tiles of e->tiles)

struct GlobalAcc {
    DevState S;
    const DevTopo &T;
    __device__ __forceinline__ int mon(int j) const { return T.mon_type[j]; }
    __device__ __forceinline__ bool gtp1(int j) const { return T.gtp[j] == 1; }
    __device__ __forceinline__ bool offtub(int j) const { return T.on_tub[j] == 0; }
    __device__ __forceinline__ float4 xyz(int j) const { return S.xyz[j]; }
    __device__ __forceinline__ float4 ang(int j) const { return S.ang[j]; }
    __device__ __forceinline__ float4 fn(int j) const { return S.fn[j]; }
    __device__ __forceinline__ float4 fp1(int j) const { return S.fp1[j]; }
    __device__ __forceinline__ float4 fp2(int j) const { return S.fp2[j]; }
};

__global__ void __launch_bounds__(STREAM_BLOCK) k_frames(DevState S, int nb, LatOffsets L) {
    const int b = blockIdx.x * STREAM_BLOCK + threadIdx.x;
    if (b >= nb) return;
    const float4 a = S.ang[b];
    float4 n, p1, p2;
    frame_vectors(a.x, a.y, a.z, L, n, p1, p2);
    S.fn[b] = n;
    S.fp1[b] = p1;
    S.fp2[b] = p2;
}

// Force + Euler-Maruyama step.  Reads Sin, writes Sout (both positions and
// rotated offsets), advances the HybridTaus states of mobile beads only
// (src/compute_cuda.cu:949-953) unless noise is injected.
template <int BS>
__global__ void __launch_bounds__(BS)
    k_step_plain(mte_params P, DevTopo T, DevState Sin, DevState Sout, uint4 *seeds, const float *noise,
                 LatOffsets L) {
    const int b = xcd_bead(BS);
    if (b >= T.nb) return;
    float4 x = Sin.xyz[b], a = Sin.ang[b];
    const bool extra = T.extra[b];
    const bool mobile = !T.fixed[b] && !extra;
    if (!mobile) {
        Sout.xyz[b] = x;
        Sout.ang[b] = a;
        Sout.fn[b] = Sin.fn[b];
        Sout.fp1[b] = Sin.fp1[b];
        Sout.fp2[b] = Sin.fp2[b];
        return;
    }
    Own I = make_own(x, a, Sin.fn[b], Sin.fp1[b], Sin.fp2[b]);
    set_box(I, T.tbox[T.tid[b]]);
    float F[6];
    bead_force(P, GlobalAcc{Sin, T}, GlobalLists{T, b, 0}, b, P.is_wall ? T.zs[T.tid[b]] : 0.f, I, F);
    float4 rx, ra;
    if (noise) {
        const float *nz = noise + 6 * (size_t)b;
        rx = make_float4(nz[0], nz[1], nz[2], 0.f);
        ra = make_float4(nz[3], nz[4], nz[5], 0.f);
    } else {
        uint4 s = seeds[b];
        rx = normals4(s);
        seeds[b] = s;
        s = seeds[b + T.nb];
        ra = normals4(s);
        seeds[b + T.nb] = s;
    }
    em_update(P, F, rx, ra, x, a);
    float4 n, p1, p2;
    frame_vectors(a.x, a.y, a.z, L, n, p1, p2);
    Sout.xyz[b] = x;
    Sout.ang[b] = a;
    Sout.fn[b] = n;
    Sout.fp1[b] = p1;
    Sout.fp2[b] = p2;
}

// Random-force draw for TEA steps, fused into the force pass: every bead
// draws sqrt(2 kB T gammaR / dt) xi from its xyz stream
// (integrateTea_prepare, src/bdhitea_kernel.cu:16-36).
struct TeaDraw {
    uint4 *seeds;
    const float *noise;
    float4 *R;   // nullptr: no draw
};

// Forces (compute_kernel semantics; extra beads get zero), optionally with
// the TEA random forces.
template <int BS>
__global__ void __launch_bounds__(BS)
    k_forces(mte_params P, DevTopo T, DevState S, float4 *fxyz, float4 *fang, TeaDraw D) {
    const int b = xcd_bead(BS);
    if (b >= T.nb) return;
    float F[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (!T.extra[b]) {
        Own I = make_own(S.xyz[b], S.ang[b], S.fn[b], S.fp1[b], S.fp2[b]);
        set_box(I, T.tbox[T.tid[b]]);
        bead_force(P, GlobalAcc{S, T}, GlobalLists{T, b, 0}, b, P.is_wall ? T.zs[T.tid[b]] : 0.f, I, F);
    }
    fxyz[b] = make_float4(F[0], F[1], F[2], 0.f);
    fang[b] = make_float4(F[3], F[4], F[5], 0.f);   // fi, theta, psi
    if (D.R) {
        const float var = sqrtf(2.0f * KB * P.temp * P.gammaR / P.dt);
        float4 r;
        if (D.noise) {
            const float *nz = D.noise + 6 * (size_t)b;
            r = make_float4(nz[0], nz[1], nz[2], 0.f);
        } else {
            uint4 s = D.seeds[b];
            r = normals4(s);
            D.seeds[b] = s;
        }
        D.R[b] = make_float4(r.x * var, r.y * var, r.z * var, 0.f);
    }
}

// The same forces with the bead's whole trajectory staged in LDS (positions,
// rotated offsets, angles and flags of every bead), one workgroup per
// 256-bead tile of a trajectory: neighbour data are read from LDS instead of
// gathered from L2 per pair.  For trajectories that fit (forces_lds_bytes).
__global__ void __launch_bounds__(FORCE_TILE)
    k_forces_lds(mte_params P, DevTopo T, DevState S, const int2 *tiles, float4 *fxyz, float4 *fang, TeaDraw D) {
    extern __shared__ float4 lds[];
    const int2 tl = tiles[blockIdx.x];
    const int t = tl.x, base = T.tbase[t], n = T.tsize[t];
    float4 *lx = lds, *ln = lds + n, *l1 = ln + n, *l2 = l1 + n;
    stage_state(T, S, base, n, lx, ln, l1, l2);
    __syncthreads();
    const int i = tl.y + threadIdx.x;
    if (i >= n) return;
    const int b = base + i;
    float F[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (!T.extra[b]) {
        const float4 fn = ln[i], f1 = l1[i], f2 = l2[i];
        Own I = make_own(lx[i], make_float4(fn.w, f1.w, f2.w, 0.f), fn, f1, f2);
        set_box(I, T.tbox[t]);
        bead_force(P, LdsAcc{lx, ln, l1, l2}, GlobalLists{T, b, base}, i, P.is_wall ? T.zs[t] : 0.f, I, F);
    }
    fxyz[b] = make_float4(F[0], F[1], F[2], 0.f);
    fang[b] = make_float4(F[3], F[4], F[5], 0.f);   // fi, theta, psi
    if (D.R) {
        const float var = sqrtf(2.0f * KB * P.temp * P.gammaR / P.dt);
        float4 r;
        if (D.noise) {
            const float *nz = D.noise + 6 * (size_t)b;
            r = make_float4(nz[0], nz[1], nz[2], 0.f);
        } else {
            uint4 s = D.seeds[b];
            r = normals4(s);
            D.seeds[b] = s;
        }
        D.R[b] = make_float4(r.x * var, r.y * var, r.z * var, 0.f);
    }
}

// Dynamic LDS of k_forces_lds for trajectories of up to nmax beads.
size_t forces_lds_bytes(int nmax) { return (size_t)4 * sizeof(float4) * nmax; }

template <int BS>
__global__ void __launch_bounds__(BS)
    k_energies(mte_params P, DevTopo T, DevState S, double *en) {
    const int b = xcd_bead(BS);
    if (b >= T.nb) return;
    float U[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (!T.extra[b]) {
        Own I = make_own(S.xyz[b], S.ang[b], S.fn[b], S.fp1[b], S.fp2[b]);
        set_box(I, T.tbox[T.tid[b]]);
        bead_energy(P, T, GlobalAcc{S, T}, b, I, U);
    }
#pragma unroll
    for (int q = 0; q < 7; q++) en[7 * (size_t)b + q] = (double)U[q];
}

// Per-trajectory FP64 sums of the per-bead energies (OutputAllEnergies,
// src/updater.cpp:3-43), one workgroup per trajectory, fixed reduction order.
__global__ void __launch_bounds__(256) k_traj_sum(DevTopo T, const double *en, double *out) {
    __shared__ double red[7][256];
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += 256)
        for (int q = 0; q < 7; q++) acc[q] += en[7 * (size_t)(base + i) + q];
    for (int q = 0; q < 7; q++) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 7; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 7) out[7 * t + threadIdx.x] = red[threadIdx.x][0];
}

static inline int grid_for(int n, int bs = STREAM_BLOCK) { return (n + bs - 1) / bs; }
// Grid of a kernel that takes its beads through xcd_bead: a multiple of 8.
static inline int grid_xcd(int n, int bs) { return 8 * ((grid_for(n, bs) + 7) / 8); }

// Launch of a per-bead kernel template at the tuned workgroup size.
#define MTE_LAUNCH_BS(KERNEL, e, ...)                                                                              \
    do {                                                                                                           \
        const int bs_ = (e)->tune.stream_block;                                                                    \
        if (bs_ == 64)                                                                                             \
            hipLaunchKernelGGL(KERNEL<64>, dim3(grid_xcd((e)->nb, 64)), dim3(64), 0, (e)->stream, __VA_ARGS__);   \
        else if (bs_ == 128)                                                                                       \
            hipLaunchKernelGGL(KERNEL<128>, dim3(grid_xcd((e)->nb, 128)), dim3(128), 0, (e)->stream, __VA_ARGS__); \
        else                                                                                                       \
            hipLaunchKernelGGL(KERNEL<256>, dim3(grid_xcd((e)->nb, 256)), dim3(256), 0, (e)->stream, __VA_ARGS__); \
    } while (0)

// The updater's tubule classification (mt_length, src/updater.cpp:164-205)
// on the device: a bead is on the tubule when rad_lo < sqrt(x^2 + y^2) <
// rad_hi and cos(theta) > cos_min.  Writes the per-bead flag, optionally
// into the barrier's on_tubule array too, and counts per trajectory.
__global__ void __launch_bounds__(STREAM_BLOCK)
    k_mt_length(DevTopo T, DevState S, float rad_hi, float rad_lo, float cos_min, int apply, int *flags,
                int *counts) {
    const int b = blockIdx.x * STREAM_BLOCK + threadIdx.x;
    if (b >= T.nb) return;
    const float4 x = S.xyz[b];
    const float rad = sqrtf(x.x * x.x + x.y * x.y);
    const int on = (rad < rad_hi && rad > rad_lo && cosf(S.ang[b].y) > cos_min) ? 1 : 0;
    flags[b] = on;
    if (apply) T.on_tub[b] = on;
    if (on) atomicAdd(&counts[T.tid[b]], 1);
}

__global__ void __launch_bounds__(256) k_zero_ints(int *p, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_mt_length(mte_engine *e, float rad_hi, float rad_lo, float cos_min, int apply) {
    hipLaunchKernelGGL(k_zero_ints, dim3((e->ntraj + 255) / 256), dim3(256), 0, e->stream, e->mtl_counts, e->ntraj);
    hipLaunchKernelGGL(k_mt_length, dim3(grid_for(e->nb)), dim3(STREAM_BLOCK), 0, e->stream, e->T, e->S[e->cur],
                       rad_hi, rad_lo, cos_min, apply, e->mtl_flags, e->mtl_counts);
    return hipGetLastError();
}

hipError_t launch_frames(mte_engine *e, int buf) {
    hipLaunchKernelGGL(k_frames, dim3(grid_for(e->nb)), dim3(STREAM_BLOCK), 0, e->stream, e->S[buf], e->nb, e->L);
    return hipGetLastError();
}

hipError_t launch_step_plain(mte_engine *e, const float *noise_step) {
    MTE_LAUNCH_BS(k_step_plain, e, e->P, e->T, e->S[e->cur], e->S[1 - e->cur], e->seeds, noise_step, e->L);
    e->cur = 1 - e->cur;
    return hipGetLastError();
}

static hipError_t launch_forces_draw(mte_engine *e, TeaDraw D) {
    if (e->tune.forces == MTE_FORCES_LDS && e->forces_lds > 0) {
        const hipError_t s = hipFuncSetAttribute((const void *)k_forces_lds,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->forces_lds);
        if (s != hipSuccess) return s;
        hipLaunchKernelGGL(k_forces_lds, dim3(e->ntiles), dim3(FORCE_TILE), e->forces_lds, e->stream, e->P, e->T,
                           e->S[e->cur], e->tiles, e->fxyz, e->fang, D);
    } else {
        MTE_LAUNCH_BS(k_forces, e, e->P, e->T, e->S[e->cur], e->fxyz, e->fang, D);
    }
    return hipGetLastError();
}

hipError_t launch_forces(mte_engine *e) { return launch_forces_draw(e, TeaDraw{nullptr, nullptr, nullptr}); }

hipError_t launch_forces_tea(mte_engine *e, const float *noise_step) {
    return launch_forces_draw(e, TeaDraw{e->seeds, noise_step, e->tea.R});
}

hipError_t launch_energies(mte_engine *e) {
    MTE_LAUNCH_BS(k_energies, e, e->P, e->T, e->S[e->cur], e->en_bead);
    hipLaunchKernelGGL(k_traj_sum, dim3(e->ntraj), dim3(256), 0, e->stream, e->T, e->en_bead, e->en_traj);
    return hipGetLastError();
}

}  // namespace mte
