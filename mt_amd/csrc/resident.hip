// resident.hip -- plain Langevin dynamics with each trajectory resident in
// one workgroup's LDS across many steps.
//
// One workgroup owns one trajectory: positions, angles and rotated offsets of
// all its beads (80 B per bead) live in LDS, each thread owns PT beads and
// keeps their two HybridTaus states in registers.  A launch runs every step up
// to the next list-rebuild boundary (LJPairsUpdateFreq steps, 100 in the
// batched workload), doing the rebuild itself first when it is due, so HBM
// sees the state once per launch instead of twice per step.  Per step:
//   forces + Euler-Maruyama for own beads (neighbours read from LDS, lists
//   from L2-resident k-major arrays) -> barrier -> write own new state and
//   rotated offsets to LDS -> barrier.
// The arithmetic is bead_force / em_update, the same code the streamed step
// runs, so the two modes agree to rounding.
#include "bead_physics.h"
#include "engine_internal.h"
#include "rebuild_common.h"

namespace mte {

constexpr int RES_MAX_THREADS = 1024;  // one bead per thread
constexpr int RES_MULTI_THREADS = 768;  // several beads per thread: leaves room for more VGPRs
constexpr int RES_MAX_PT = 3;
constexpr int LDS_ARRAYS = 5;   // xyz, ang, fn, fp1, fp2

// Per-bead flags packed into xyz.w in LDS.
constexpr int FL_MON = 1, FL_GTP1 = 2, FL_OFFTUB = 4;

struct LdsAcc {
    const float4 *x, *a, *n, *p1, *p2;
    int base;
    __device__ __forceinline__ float4 xyz(int j) const { return x[j - base]; }
    __device__ __forceinline__ float4 ang(int j) const { return a[j - base]; }
    __device__ __forceinline__ float4 fn(int j) const { return n[j - base]; }
    __device__ __forceinline__ float4 fp1(int j) const { return p1[j - base]; }
    __device__ __forceinline__ float4 fp2(int j) const { return p2[j - base]; }
    __device__ __forceinline__ int flags(int j) const { return __float_as_int(x[j - base].w); }
    __device__ __forceinline__ int mon(int j) const { return flags(j) & FL_MON; }
    __device__ __forceinline__ bool gtp1(int j) const { return (flags(j) & FL_GTP1) != 0; }
    __device__ __forceinline__ bool offtub(int j) const { return (flags(j) & FL_OFFTUB) != 0; }
};

template <int PT, int MAXT>
__global__ void __launch_bounds__(MAXT)
    k_resident(mte_params P, DevTopo T, DevState S, uint4 *seeds, const float *noise, int nsteps, int rebuild,
               int do_lj, int do_bonded, LatOffsets L, int stride) {
    extern __shared__ float4 lds[];
    float4 *lx = lds, *la = lds + stride, *ln = lds + 2 * stride, *l1 = lds + 3 * stride, *l2 = lds + 4 * stride;
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    const int nb = T.nb;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = base + i;
        float4 x = S.xyz[b];
        const int fl = (T.mon_type[b] ? FL_MON : 0) | (T.gtp[b] == 1 ? FL_GTP1 : 0) | (T.on_tub[b] == 0 ? FL_OFFTUB : 0);
        x.w = __int_as_float(fl);
        lx[i] = x;
        la[i] = S.ang[b];
        ln[i] = S.fn[b];
        l1[i] = S.fp1[b];
        l2[i] = S.fp2[b];
    }
    bool mob[PT];
    uint4 s0[PT], s1[PT];
#pragma unroll
    for (int k = 0; k < PT; k++) {
        const int i = threadIdx.x + k * blockDim.x;
        const int b = base + i;
        mob[k] = i < n && !T.fixed[b] && !T.extra[b];
        if (mob[k] && !noise) {
            s0[k] = seeds[b];
            s1[k] = seeds[b + nb];
        }
    }
    __syncthreads();
    const LdsAcc A{lx, la, ln, l1, l2, base};
    if (rebuild) {
#pragma unroll
        for (int k = 0; k < PT; k++) {
            const int i = threadIdx.x + k * blockDim.x;
            if (i >= n) continue;
            const int b = base + i;
            RebuildOwn o = rebuild_own(lx[i], ln[i], l1[i], l2[i], A.mon(b), T.harm_raw[b], i, b, base);
            if (!T.extra[b])
                for (int j = 0; j < n; j++)
                    rebuild_pair(T, o, j, lx[j], ln[j], l1[j], l2[j], __float_as_int(lx[j].w) & FL_MON,
                                 P.ljpairscutoff, do_lj, do_bonded);
            rebuild_finish(T, o, do_lj, do_bonded);
        }
    }
    for (int s = 0; s < nsteps; s++) {
        float4 nx[PT], na[PT];
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (!mob[k]) continue;
            const int i = threadIdx.x + k * blockDim.x;
            const int b = base + i;
            float4 x = lx[i], a = la[i];
            const Own I = make_own(x, a, ln[i], l1[i], l2[i]);
            float F[6];
            bead_force(P, T, A, b, I, F);
            float4 rx, ra;
            if (noise) {
                const float *nz = noise + (size_t)s * 6 * nb + 6 * (size_t)b;
                rx = make_float4(nz[0], nz[1], nz[2], 0.f);
                ra = make_float4(nz[3], nz[4], nz[5], 0.f);
            } else {
                rx = normals4(s0[k]);
                ra = normals4(s1[k]);
            }
            em_update(P, F, rx, ra, x, a);
            nx[k] = x;
            na[k] = a;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PT; k++) {
            if (!mob[k]) continue;
            const int i = threadIdx.x + k * blockDim.x;
            lx[i] = nx[k];   // w keeps the packed flags
            la[i] = na[k];
            float4 nn, p1, p2;
            frame_vectors(na[k].x, na[k].y, na[k].z, L, nn, p1, p2);
            ln[i] = nn;
            l1[i] = p1;
            l2[i] = p2;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = base + i;
        float4 x = lx[i];
        x.w = 0.f;
        S.xyz[b] = x;
        S.ang[b] = la[i];
        S.fn[b] = ln[i];
        S.fp1[b] = l1[i];
        S.fp2[b] = l2[i];
    }
#pragma unroll
    for (int k = 0; k < PT; k++) {
        if (mob[k] && !noise) {
            const int b = base + threadIdx.x + k * blockDim.x;
            seeds[b] = s0[k];
            seeds[b + nb] = s1[k];
        }
    }
}

bool resident_fits(const mte_engine *e, ResidentPlan *plan) {
    *plan = ResidentPlan{};
    if (e->P.tea_on) return false;
    int nmax = 0;
    for (int t = 0; t < e->ntraj; t++) nmax = std::max<int>(nmax, (int)(e->off[t + 1] - e->off[t]));
    int lds_limit = 0;
    if (hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) != hipSuccess)
        return false;
    const size_t bytes = (size_t)LDS_ARRAYS * sizeof(float4) * nmax;
    if (bytes > (size_t)lds_limit) return false;
    int pt = 1;
    if (nmax > RES_MAX_THREADS) pt = (nmax + RES_MULTI_THREADS - 1) / RES_MULTI_THREADS;
    if (pt > RES_MAX_PT) return false;
    int threads = (nmax + pt - 1) / pt;
    threads = ((threads + 63) / 64) * 64;
    plan->threads = threads;
    plan->per_thread = pt;
    plan->max_beads = nmax;
    plan->lds_bytes = bytes;
    return true;
}

template <int PT, int MAXT>
static hipError_t launch_pt(mte_engine *e, int nsteps, int rb, int do_lj, int do_bonded, const float *noise) {
    const ResidentPlan &pl = e->rplan;
    hipError_t s = hipFuncSetAttribute((const void *)k_resident<PT, MAXT>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds_bytes);
    if (s != hipSuccess) return s;
    hipLaunchKernelGGL((k_resident<PT, MAXT>), dim3(e->ntraj), dim3(pl.threads), pl.lds_bytes, e->stream, e->P,
                       e->T, e->S[e->cur], e->seeds, noise, nsteps, rb, do_lj, do_bonded, e->L, pl.max_beads);
    return hipGetLastError();
}

hipError_t launch_resident(mte_engine *e, int nsteps, bool rebuild_first, const float *noise) {
    const ResidentPlan &pl = e->rplan;
    const int do_lj = e->P.lj_on ? 1 : 0;
    const int do_bonded = e->P.is_assembly ? 1 : 0;
    const int rb = (rebuild_first && (do_lj || do_bonded)) ? 1 : 0;
    switch (pl.per_thread) {
    case 1: return launch_pt<1, RES_MAX_THREADS>(e, nsteps, rb, do_lj, do_bonded, noise);
    case 2: return launch_pt<2, RES_MULTI_THREADS>(e, nsteps, rb, do_lj, do_bonded, noise);
    case 3: return launch_pt<3, RES_MULTI_THREADS>(e, nsteps, rb, do_lj, do_bonded, noise);
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mte
