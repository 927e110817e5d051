// tea.hip -- TEA hydrodynamic interactions (Geyer & Winter 2009) on gfx950.
//
// Restates src/bdhitea_kernel.cu:16-213 and the host FP64 normalisation of
// src/bdhitea.cu:57-118, with the normalisation moved on device in FP64:
//   per bead  ci_i = sum_j rowwise (D_hat_ij)^2, eps_i = sum_j sum(D_hat_ij)
//   per traj  eps = sum_i eps_i / (3N(3N-3)), a = (3N-1)eps^2 - (3N-2)eps,
//             beta = (1 - sqrt(1-a))/a  (0.5 when |a| < 1e-7)
//   per bead  C_i = 1/sqrt(1 + beta^2 ci_i),
//             f_i = F_i + C_i.R_i + sum_{j!=i} D_hat_ij (F_j + beta C_i.R_j),
//             r_i += dt/gammaR f_i
// with D_hat_ij = D_ij/D_ii the Rotne-Prager-Yamakawa block, N the count of
// non-extra beads of the trajectory.
//
// Work layout of the per-step product: work units of (trajectory, 256-bead i
// group, j chunk of 64, 128 or 256 beads as tuned), the chunk's positions, forces and random forces in
// LDS, FP32 partial sums per chunk reduced in fixed order by the finish kernel,
// which adds the self term and integrates.  Each (i, j) block costs one rsqrt
// and ~35 FP32 operations: the product is VALU-bound.
#include "engine_internal.h"
#include <type_traits>

namespace mte {

constexpr int TEA_TILE = 256;

struct RpyCoef { float crr, cii; };

// RPY coefficients of D_hat = cii I + crr e e^T at r^2 (rinv = 1/r), both
// branches of src/bdhitea_kernel.cu:38-58 evaluated and selected (no
// divergence); ia = 1/a.
__device__ __forceinline__ RpyCoef rpy_select(float r2, float rinv, float a, float ia) {
    const float u1 = a * rinv, u2 = u1 * u1;
    const float ra = r2 * rinv * ia;
    const bool near = !(ra > 2.f);
    RpyCoef c;
    c.cii = near ? fmaf(-0.28125f, ra, 1.f) : 0.75f * u1 * fmaf(0.66666666666666667f, u2, 1.f);
    c.crr = near ? 0.09375f * ra : 0.75f * u1 * fmaf(-2.f, u2, 1.f);
    return c;
}

// Lane l of the wave receives lane (l + 1) mod 64's value (all lanes active).
__device__ __forceinline__ float wave_rol1(float v) {
    const int x = __float_as_int(v);
    return __int_as_float(__builtin_amdgcn_update_dpp(x, x, 0x134, 0xF, 0xF, false));
}

// wave_rol1(v) + p, the rotation a DPP operand of the addition.
__device__ __forceinline__ float wave_rol1_add(float v, float p) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, true)) + p;
}

// The far branch of rpy_select alone (r > 2a).
__device__ __forceinline__ RpyCoef rpy_far(float rinv, float a) {
    const float u1 = a * rinv, u2 = u1 * u1;
    RpyCoef c;
    c.cii = 0.75f * u1 * fmaf(0.66666666666666667f, u2, 1.f);
    c.crr = 0.75f * u1 * fmaf(-2.f, u2, 1.f);
    return c;
}

// Epsilon partials and ci of bead i over all j of its trajectory
// (integrateTea_epsilon_unlisted, src/bdhitea_kernel.cu:61-101), with the
// block sums in closed form (|e| = 1):
//   sum of the 9 entries of D_hat      = 3 cii + crr (ex + ey + ez)^2
//   row a sum of squares               = cii^2 + crr (2 cii + crr) e_a^2
// Extra beads enter with weight 0; the j == i slot (d = 0, its distance
// clamped finite) is taken out after the loop, so no branch sits in the pair
// loop.  PERIODIC: some trajectory has a periodic box.
template <bool PERIODIC>
__global__ void __launch_bounds__(TEA_TILE)
    k_tea_eps(mte_params P, DevTopo T, DevState S, const int2 *tiles, float4 *ci, float *eps) {
    __shared__ float4 sx[TEA_TILE];
    const int2 tl = tiles[blockIdx.x];
    const int base = T.tbase[tl.x], n = T.tsize[tl.x];
    const int li = tl.y + threadIdx.x;
    const bool active = li < n;
    const int b = base + li;
    float3 ri = make_float3(0, 0, 0);
    bool skip = true;
    if (active) {
        ri = f3(S.xyz[b]);
        skip = T.extra[b] != 0;
    }
    const float box = T.tbox[tl.x], ibox = inv_side(box);
    const float a = P.tea_a, ia = rcp_fast(P.tea_a);
    float scc = 0.f, sx2 = 0.f, sy2 = 0.f, sz2 = 0.f, sall = 0.f;
    for (int j0 = 0; j0 < n; j0 += TEA_TILE) {
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            float4 v = S.xyz[base + jl];
            v.w = T.extra[base + jl] ? 0.f : 1.f;
            sx[threadIdx.x] = v;
        }
        __syncthreads();
        const int m = min(TEA_TILE, n - j0);
        if (active && !skip) {
            // not unrolled: measured faster than unrolled by 2 or 4
#pragma unroll 1
            for (int jj = 0; jj < m; jj++) {
                const float4 xj = sx[jj];
                float3 d = sub3(f3(xj), ri);
                if (PERIODIC) d = min_image_inv(d, box, ibox);
                const float wj = xj.w;
                const float r2 = fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, 1e-12f);
                const float rinv = rsq_fast(r2);
                const RpyCoef c = rpy_select(r2, rinv, a, ia);
                const float ex = d.x * rinv, ey = d.y * rinv, ez = d.z * rinv;
                const float cw = wj * c.cii;
                const float k = wj * c.crr * fmaf(2.f, c.cii, c.crr);
                const float se = ex + ey + ez;
                scc = fmaf(cw, c.cii, scc);
                sx2 = fmaf(k, ex * ex, sx2);
                sy2 = fmaf(k, ey * ey, sy2);
                sz2 = fmaf(k, ez * ez, sz2);
                sall += fmaf(wj * c.crr, se * se, 3.f * cw);
            }
        }
        __syncthreads();
    }
    if (active && !skip) {
        // the j == i slot entered as a coincident bead of weight 1 (d = 0):
        // cii^2 into scc and 3 cii into sall, recomputed and taken out
        const float r2 = 1e-12f, rinv = rsq_fast(r2);
        const float cii = rpy_select(r2, rinv, a, ia).cii;
        scc = fmaf(-cii, cii, scc);
        sall = fmaf(-3.f, cii, sall);
    }
    if (active) {
        ci[b] = make_float4(scc + sx2, scc + sy2, scc + sz2, 0.f);
        eps[b] = sall;
    }
}

// Epsilon partials and ci with each pair evaluated once.  The five sums of
// k_tea_eps are even in e (cii^2, crr (2 cii + crr) e_a^2, 3 cii + crr
// (ex + ey + ez)^2), so a pair adds the same terms to both beads.  Work unit:
// (trajectory, 1024-bead group I, group J >= I), one workgroup of
// EPS_WAVES waves; wave w takes chunks c = 16 I + w + h EPS_WAVES, h <
// EPS_PER (64 beads each, one bead of each per lane) against each 64-bead
// chunk c_J of group J in turn, staged in LDS:
//   c_J > c    both sides, lane l meeting j = (l + s) mod 64 at step s, the
//              j-side sums carried along by a one-lane rotation (wave_rol1);
//              when c_J lies above all of the wave's chunks, the lane's
//              beads meet the same j together and share the rotation,
//   c_J = c    one-sided over the chunk, the j == i term taken out after,
//   c_J < c    nothing (that pair of chunks is the wave c_J's).
// The waves' j-side sums of a chunk are added in wave order in LDS.  Output
// (deterministic, no atomics): i-side sums of a bead of group H into
// part[0][J][.] for J >= H, j-side sums into part[1][I][.] for I <= H; the
// pair weight is 1 for two non-extra beads and 0 otherwise (WEIGHTED: some
// bead of the batch is extra; without, every weight is 1 and none is formed).
constexpr int EPS_CHUNKS = EPS_GROUP / 64;        // 64-bead chunks of a group
constexpr int EPS_PER = 4;                        // chunks of a wave
constexpr int EPS_WAVES = EPS_CHUNKS / EPS_PER;   // waves of a workgroup
static_assert(EPS_CHUNKS % EPS_PER == 0 && EPS_WAVES >= 2, "chunks split over the waves; 128 threads stage c_J");
template <bool PERIODIC, bool WEIGHTED>
__global__ void __launch_bounds__(64 * EPS_WAVES)
    k_tea_eps_sym(mte_params P, DevTopo T, DevState S, const int4 *units, int ngrp, float *part) {
    __shared__ float4 jx[2 * 64];              // chunk c_J: x, weight; twice, no wrap
    __shared__ float jred[EPS_WAVES][5][64];   // j-side sums of each wave
    const int4 u = units[blockIdx.x];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int base = T.tbase[u.x], n = T.tsize[u.x], nb = T.nb;
    const int nch = (n + 63) / 64;
    // Lengths in the pair loop are in box units in a periodic box (positions
    // x / box, so the minimum image is d - rint(d)) and in nm otherwise:
    // sc is the unit's inverse, wf 1 where the image is taken.
    const float box = T.tbox[u.x];
    const bool wrap = PERIODIC && box > 0.f;
    const float sc = wrap ? inv_side(box) : 1.f, wf = wrap ? 1.f : 0.f;
    const float a = P.tea_a * sc, ia = rcp_fast(P.tea_a) * (wrap ? box : 1.f);
    const float r2min = 1e-12f * sc * sc;
    // the lane's bead in each of the wave's chunks
    int cI[EPS_PER], li[EPS_PER];
    bool vi[EPS_PER];
    float3 xi[EPS_PER];
    float wi[EPS_PER], ai[EPS_PER][5];
#pragma unroll
    for (int h = 0; h < EPS_PER; h++) {
        cI[h] = u.y * EPS_CHUNKS + w + h * EPS_WAVES;
        li[h] = cI[h] * 64 + l;
        vi[h] = li[h] < n;
        xi[h] = make_float3(0.f, 0.f, 0.f);
        wi[h] = 0.f;
        if (vi[h]) {
            const float3 x = f3(S.xyz[base + li[h]]);
            xi[h] = make_float3(x.x * sc, x.y * sc, x.z * sc);
            wi[h] = T.extra[base + li[h]] ? 0.f : 1.f;
        }
#pragma unroll
        for (int k = 0; k < 5; k++) ai[h][k] = 0.f;
    }
    const int c0 = u.z * EPS_CHUNKS, c1 = min(c0 + EPS_CHUNKS, nch);
    // one pair's five terms p: added to the sums t; with t2 (both sides),
    // t2 += p, after t2 takes the next lane's t2 when ROT (the j-side
    // rotation, done with the first of the lane's beads to meet this j).  A
    // wave with no pair inside 2a (margin over the rounding of r / a) keeps
    // the far coefficients, which is the branch rpy_select picks for each of
    // its pairs.  WT: the pair weights are formed (extra beads in the batch,
    // or the padding of a trajectory's last chunk, whose slots have weight 0).
    const float near2 = (2.002f * a) * (2.002f * a);
    auto pair = [&](auto wt, auto rot, const float3 &xa, float wa, float4 xj, float *t, float *t2) {
        constexpr bool WT = decltype(wt)::value, ROT = decltype(rot)::value;
        float3 d = make_float3(xj.x - xa.x, xj.y - xa.y, xj.z - xa.z);
        if (PERIODIC) {
            d.x = fmaf(-wf, rintf(d.x), d.x);
            d.y = fmaf(-wf, rintf(d.y), d.y);
            d.z = fmaf(-wf, rintf(d.z), d.z);
        }
        const float r2 = fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, r2min);
        const float rinv = rsq_fast(r2);
        RpyCoef c = rpy_far(rinv, a);
        if (__any(r2 <= near2)) {
            __asm__ volatile("");   // a branch: no speculation of the near form
            const float ra = r2 * rinv * ia;
            const bool nr = !(ra > 2.f);
            c.cii = nr ? fmaf(-0.28125f, ra, 1.f) : c.cii;
            c.crr = nr ? 0.09375f * ra : c.crr;
        }
        const float ww = WT ? wa * xj.w : 1.f;
        const float ex = d.x * rinv, ey = d.y * rinv, ez = d.z * rinv;
        const float cw = WT ? ww * c.cii : c.cii;
        const float wcrr = WT ? ww * c.crr : c.crr;
        const float k = wcrr * fmaf(2.f, c.cii, c.crr);
        const float se = ex + ey + ez;
        const float tall = fmaf(wcrr, se * se, 3.f * cw);
        const float qx = ex * ex, qy = ey * ey, qz = ez * ez;
        t[0] = fmaf(cw, c.cii, t[0]);
        t[1] = fmaf(k, qx, t[1]);
        t[2] = fmaf(k, qy, t[2]);
        t[3] = fmaf(k, qz, t[3]);
        t[4] += tall;
        if (t2) {
            t2[0] = fmaf(cw, c.cii, ROT ? wave_rol1(t2[0]) : t2[0]);
            t2[1] = fmaf(k, qx, ROT ? wave_rol1(t2[1]) : t2[1]);
            t2[2] = fmaf(k, qy, ROT ? wave_rol1(t2[2]) : t2[2]);
            t2[3] = fmaf(k, qz, ROT ? wave_rol1(t2[3]) : t2[3]);
            t2[4] = ROT ? wave_rol1_add(t2[4], tall) : t2[4] + tall;
        }
    };
    const std::true_type yes{};
    const std::false_type no{};
    for (int cJ = c0; cJ < c1; cJ++) {
        if (threadIdx.x < 128) {
            const int lj = cJ * 64 + (threadIdx.x % 64);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lj < n) {
                v = S.xyz[base + lj];
                v.x *= sc;
                v.y *= sc;
                v.z *= sc;
                v.w = T.extra[base + lj] ? 0.f : 1.f;
            }
            jx[threadIdx.x] = v;
        }
        __syncthreads();
        float gj[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        // step s of a both-sided sweep: lane l meets j = (l + s) mod 64, its
        // gj the sums of that j (rotated in from lane l + 1, which held j at
        // the sweep's step s - 1; sweeps in a row continue the rotation).
        // A chunk below c_J is full; c_J may be the padded last one.
        bool two[EPS_PER], all = true;
#pragma unroll
        for (int h = 0; h < EPS_PER; h++) {
            two[h] = cI[h] < nch && cJ > cI[h];
            all = all && two[h];
        }
        const bool wtj = WEIGHTED || (cJ + 1) * 64 > n;
        if (all) {
            if (wtj)
                for (int s = 0; s < 64; s++) {
                    const float4 xj = jx[l + s];
                    pair(yes, yes, xi[0], wi[0], xj, ai[0], gj);
#pragma unroll
                    for (int h = 1; h < EPS_PER; h++) pair(yes, no, xi[h], wi[h], xj, ai[h], gj);
                }
            else
                for (int s = 0; s < 64; s++) {
                    const float4 xj = jx[l + s];
                    pair(no, yes, xi[0], wi[0], xj, ai[0], gj);
#pragma unroll
                    for (int h = 1; h < EPS_PER; h++) pair(no, no, xi[h], wi[h], xj, ai[h], gj);
                }
        } else {
#pragma unroll
            for (int h = 0; h < EPS_PER; h++)
                if (two[h]) {
                    if (wtj)
                        for (int s = 0; s < 64; s++) pair(yes, yes, xi[h], wi[h], jx[l + s], ai[h], gj);
                    else
                        for (int s = 0; s < 64; s++) pair(no, yes, xi[h], wi[h], jx[l + s], ai[h], gj);
                }
        }
#pragma unroll
        for (int h = 0; h < EPS_PER; h++)
            if (cI[h] < nch && cJ == cI[h]) {
                // one-sided over the chunk's beads: no padding slot is met
                const int jn = min(64, n - cJ * 64);
                for (int q = 0; q < jn; q++)
                    pair(std::integral_constant<bool, WEIGHTED>{}, no, xi[h], wi[h], jx[q], ai[h], nullptr);
                if (!WEIGHTED || wi[h] != 0.f) {
                    // the j == i slot entered as a coincident bead (d = 0):
                    // cii^2 into the first sum and 3 cii into the last, taken out
                    const float r2 = r2min, rinv = rsq_fast(r2);
                    const float cii = rpy_select(r2, rinv, a, ia).cii;
                    ai[h][0] = fmaf(-cii, cii, ai[h][0]);
                    ai[h][4] = fmaf(-3.f, cii, ai[h][4]);
                }
            }
        // after a multiple of 64 rotations lane l holds the sums of j =
        // (l + 63) mod 64; with none, gj is zero
#pragma unroll
        for (int k = 0; k < 5; k++) jred[w][k][(l + 63) & 63] = gj[k];
        __syncthreads();
        for (int e = threadIdx.x; e < 5 * 64; e += 64 * EPS_WAVES) {
            const int k = e / 64, q = e % 64;
            const int lj = cJ * 64 + q;
            if (lj < n) {
                float sum = 0.f;
                for (int v = 0; v < EPS_WAVES; v++) sum += jred[v][k][q];
                part[(((size_t)1 * ngrp + u.y) * 5 + k) * nb + base + lj] = sum;
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < EPS_PER; h++)
        if (vi[h]) {
#pragma unroll
            for (int k = 0; k < 5; k++) part[(((size_t)0 * ngrp + u.z) * 5 + k) * nb + base + li[h]] = ai[h][k];
        }
}

// Per bead: the group sums of k_tea_eps_sym added in fixed order into ci and
// eps, as k_tea_eps writes them.
__global__ void __launch_bounds__(256)
    k_tea_eps_reduce(DevTopo T, int ngrp, const float *part, float4 *ci, float *eps) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= T.nb) return;
    const int t = T.tid[b], n = T.tsize[t];
    const int h = (b - T.tbase[t]) / EPS_GROUP, g = (n + EPS_GROUP - 1) / EPS_GROUP;
    const int nb = T.nb;
    // each of the five sums in the same order (groups h..g-1 of side 0, then
    // 0..h of side 1), the five advancing together so their loads overlap
    float s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    for (int j = h; j < g; j++) {
        float v[5];
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = part[(((size_t)0 * ngrp + j) * 5 + k) * nb + b];
#pragma unroll
        for (int k = 0; k < 5; k++) s[k] += v[k];
    }
    for (int i = 0; i <= h; i++) {
        float v[5];
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = part[(((size_t)1 * ngrp + i) * 5 + k) * nb + b];
#pragma unroll
        for (int k = 0; k < 5; k++) s[k] += v[k];
    }
    ci[b] = make_float4(s[0] + s[1], s[0] + s[2], s[0] + s[3], 0.f);
    eps[b] = s[4];
}

// FP64 epsilon / beta per trajectory, one workgroup each.
__global__ void __launch_bounds__(256)
    k_tea_beta(mte_params P, DevTopo T, const float *eps, float *beta, double *stats, int *status) {
    __shared__ double red[256];
    __shared__ int cnt[256];
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    double s = 0.0;
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        s += (double)eps[base + i];
        c += T.extra[base + i] ? 0 : 1;
    }
    red[threadIdx.x] = s;
    cnt[threadIdx.x] = c;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            red[threadIdx.x] += red[threadIdx.x + k];
            cnt[threadIdx.x] += cnt[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double nn = (double)cnt[0];
        double e = red[0] / (3. * nn * (3. * nn - 3.));
        int fail = 0;
        if (e > 1.0) {
            if (P.tea_capricious) fail = 1;
            e = 1.0;
        }
        if (e > (double)P.tea_epsmax) fail = 1;
        const double a = (3. * nn - 1.) * e * e - (3. * nn - 2.) * e;
        double bt;
        if (fabs(a) < 1e-7) {
            bt = 0.5;
            if (P.tea_capricious && P.tea_a > 0.0f) fail = 1;
        } else {
            bt = (1. - sqrt(1. - a)) / a;
        }
        beta[t] = (float)bt;
        stats[2 * t] = e;
        stats[2 * t + 1] = (double)(float)bt;
        if (fail) atomicCAS(status, 0, 1 + t);
    }
}

// Pair sums of one (trajectory, 256-bead i group, j chunk) work unit:
//   partial_i = sum_{j in chunk, j != i, non-extra} D_hat_ij (F_j + beta C_i . R_j)
// The chunk's j data sit in LDS; extra beads enter with zero force and
// noise, so they contribute nothing and need no branch.
// Splitting j into chunks gives ~10 work units per SIMD at 64 x 1040 beads,
// against ~1 per SIMD for one thread per bead over the whole trajectory.
template <int CHUNK, bool PERIODIC>
__global__ void __launch_bounds__(TEA_TILE)
    k_tea_pairs(mte_params P, DevTopo T, DevState S, const float4 *fxyz, const float4 *R, const float4 *ci,
                const float *beta, const int4 *units, int chunk_stride, float4 *partial) {
    __shared__ float4 sx[CHUNK], sf[CHUNK], sr[CHUNK];
    const int4 u = units[blockIdx.x];   // traj, i0, j0, chunk index
    const int base = T.tbase[u.x], n = T.tsize[u.x];
    const int jn = min(CHUNK, n - u.z);
    for (int q = threadIdx.x; q < jn; q += TEA_TILE) {
        const int jb = base + u.z + q;
        float4 x = S.xyz[jb];
        float4 f = fxyz[jb], r = R[jb];
        if (T.extra[jb]) f = r = make_float4(0.f, 0.f, 0.f, 0.f);
        sx[q] = x;
        sf[q] = f;
        sr[q] = r;
    }
    __syncthreads();
    const int li = u.y + threadIdx.x;
    if (li >= n) return;
    const int b = base + li;
    float3 acc = make_float3(0.f, 0.f, 0.f);
    if (!T.extra[b]) {
        const float3 ri = f3(S.xyz[b]);
        const float bt = beta[u.x];
        const float4 cc = ci[b];
        const float w = bt * bt;
        const float cx = bt * rsq_fast(1.f + w * cc.x), cy = bt * rsq_fast(1.f + w * cc.y),
                    cz = bt * rsq_fast(1.f + w * cc.z);
        const float a = P.tea_a, ia = rcp_fast(P.tea_a);
        const float box = T.tbox[u.x], ibox = inv_side(box);
        for (int q = 0; q < jn; q++) {
            const float4 xj = sx[q], Fj = sf[q], Rj = sr[q];
            float3 d = sub3(f3(xj), ri);
            if (PERIODIC) d = min_image_inv(d, box, ibox);
            const float dx = d.x, dy = d.y, dz = d.z;
            // the j == i slot (d = 0) and coincident beads stay finite through the clamp
            // (extra beads carry zero F and R)
            const float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-12f);
            const float rinv = rsq_fast(r2);
            // D_hat = cii I + crr e e^T (src/bdhitea_kernel.cu:38-58)
            const RpyCoef c = rpy_select(r2, rinv, a, ia);
            const float vx = fmaf(cx, Rj.x, Fj.x), vy = fmaf(cy, Rj.y, Fj.y), vz = fmaf(cz, Rj.z, Fj.z);
            const float s = (c.crr * rinv * rinv) * (dx * vx + dy * vy + dz * vz);
            acc.x = fmaf(s, dx, fmaf(c.cii, vx, acc.x));
            acc.y = fmaf(s, dy, fmaf(c.cii, vy, acc.y));
            acc.z = fmaf(s, dz, fmaf(c.cii, vz, acc.z));
        }
        const int self = li - u.z;   // position of i in this chunk, if any
        if (self >= 0 && self < jn) {
            // The loop took the j == i slot as a coincident bead: d = 0 leaves
            // only its cii (v_i) term, recomputed here from the same operands
            // and taken out again.
            const float4 Fi = sf[self], Ri = sr[self];
            const float r2 = 1e-12f, rinv = rsq_fast(r2);
            const float cii = rpy_select(r2, rinv, a, ia).cii;
            acc.x = fmaf(-cii, fmaf(cx, Ri.x, Fi.x), acc.x);
            acc.y = fmaf(-cii, fmaf(cy, Ri.y, Fi.y), acc.y);
            acc.z = fmaf(-cii, fmaf(cz, Ri.z, Fi.z), acc.z);
        }
    }
    partial[(size_t)u.w * chunk_stride + b] = make_float4(acc.x, acc.y, acc.z, 0.f);
}

// Pair sums with each (i, j) RPY block evaluated once for both beads.  Work
// unit: (trajectory, 64-bead chunk I, 64-bead chunk J >= I), one
// single-wave workgroup each, so the dispatcher balances units wave by
// wave.  D_hat_ji = D_hat_ij, so the block serves
//   i-side  D_hat_ij (F_j + beta C_i . R_j)  into partial slot J of bead i,
//   j-side  D_hat_ij (F_i + beta C_j . R_i)  into partial slot I of bead j,
// and slot c of a bead holds the sum over chunk c, as k_tea_pairs<64> writes
// it (a unit may take two I chunks below J, I and I + 64: its lanes meet j
// with both of their beads before g moves on, slot I of a j bead holds the
// sum over both chunks and slot I + 1 zero).  Off the diagonal, lane l takes j = (l + s) mod 64 at step s: every
// lane meets every j of the chunk once and no two lanes the same j in a
// step.  The j-side sum of j travels with j in registers: after each step
// the wave rotates it one lane down (DPP wave_rol:1, lane l takes lane l+1's
// value), so at step s + 1 it sits in the lane that meets j next, and after
// 64 steps in lane j.  A diagonal unit (I == J) is one-sided, as
// k_tea_pairs.  j slots past the trajectory's end and extra beads carry zero
// F and R, so their terms vanish; their sums are not stored (extra beads
// store zero).
constexpr int SYM_CHUNK = 64;
template <bool PERIODIC>
__global__ void __launch_bounds__(SYM_CHUNK)
    k_tea_pairs_sym(mte_params P, DevTopo T, DevState S, const float4 *fxyz, const float4 *R, const float4 *ci,
                    const float *beta, const int4 *units, int nunits, int chunk_stride, float4 *partial) {
    // j data of each wave's block, packed as (x, F.x) (F.y, F.z, R.x, R.y)
    // (R.z, beta C) and written twice, at q and q + 64, so that lane l reads
    // slot l + s at step s with no wrap
    __shared__ float4 jd[3][2 * SYM_CHUNK];
    const int l = threadIdx.x;
    const int ui = blockIdx.x;
    const bool has = ui < nunits;
    const int4 u = has ? units[ui] : make_int4(0, 0, 0, 0);
    const int base = T.tbase[u.x], n = T.tsize[u.x];
    const float bt = beta[u.x], w2 = bt * bt;
    const float box = T.tbox[u.x], ibox = inv_side(box);
    // beads of this lane in chunk I and, in a two-chunk unit (u.w = 2, both
    // chunks below J), in chunk I + 64 ...
    const int ni = u.w == 2 ? 2 : 1;
    int bi[2];
    bool vi[2], ei[2];
    float3 xi[2], Fi[2], Ri[2], Ci[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int li = u.y + 64 * h + l;
        bi[h] = base + li;
        vi[h] = has && h < ni && li < n;
        ei[h] = vi[h] && T.extra[bi[h]];
        xi[h] = Fi[h] = Ri[h] = Ci[h] = make_float3(0.f, 0.f, 0.f);
        if (vi[h]) {
            xi[h] = f3(S.xyz[bi[h]]);
            if (!ei[h]) {
                Fi[h] = f3(fxyz[bi[h]]);
                Ri[h] = f3(R[bi[h]]);
                const float4 cc = ci[bi[h]];
                Ci[h] = make_float3(bt * rsq_fast(1.f + w2 * cc.x), bt * rsq_fast(1.f + w2 * cc.y),
                                    bt * rsq_fast(1.f + w2 * cc.z));
            }
        }
    }
    // ... and in chunk J, staged for the wave
    const int lj = u.z + l, bj = base + lj;
    const bool vj = has && lj < n;
    const bool ej = vj && T.extra[bj];
    {
        float3 x = make_float3(0.f, 0.f, 0.f), F = x, Rj = x, C = x;
        if (vj) {
            x = f3(S.xyz[bj]);
            if (!ej) {
                F = f3(fxyz[bj]);
                Rj = f3(R[bj]);
                const float4 cc = ci[bj];
                C = make_float3(bt * rsq_fast(1.f + w2 * cc.x), bt * rsq_fast(1.f + w2 * cc.y),
                                bt * rsq_fast(1.f + w2 * cc.z));
            }
        }
        const float4 d0 = make_float4(x.x, x.y, x.z, F.x), d1 = make_float4(F.y, F.z, Rj.x, Rj.y),
                     d2 = make_float4(Rj.z, C.x, C.y, C.z);
#pragma unroll
        for (int h = 0; h < 2; h++) {
            jd[0][l + h * SYM_CHUNK] = d0;
            jd[1][l + h * SYM_CHUNK] = d1;
            jd[2][l + h * SYM_CHUNK] = d2;
        }
    }
    __syncthreads();
    const float a = P.tea_a, ia = rcp_fast(P.tea_a);
    float3 acc[2] = {make_float3(0.f, 0.f, 0.f), make_float3(0.f, 0.f, 0.f)}, g = acc[0];
    const bool diag = u.y == u.z;
    if (has && diag) {
        if (vi[0] && !ei[0]) {
            const int jn = min(SYM_CHUNK, n - u.z);
            for (int q = 0; q < jn; q++) {
                const float4 A = jd[0][q], B = jd[1][q], Cq = jd[2][q];
                float3 d = make_float3(A.x - xi[0].x, A.y - xi[0].y, A.z - xi[0].z);
                if (PERIODIC) d = min_image_inv(d, box, ibox);
                const float r2 = fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, 1e-12f);
                const float rinv = rsq_fast(r2);
                const RpyCoef c = rpy_select(r2, rinv, a, ia);
                const float vx = fmaf(Ci[0].x, B.z, A.w), vy = fmaf(Ci[0].y, B.w, B.x), vz = fmaf(Ci[0].z, Cq.x, B.y);
                const float sv = (c.crr * rinv * rinv) * (d.x * vx + d.y * vy + d.z * vz);
                acc[0].x = fmaf(sv, d.x, fmaf(c.cii, vx, acc[0].x));
                acc[0].y = fmaf(sv, d.y, fmaf(c.cii, vy, acc[0].y));
                acc[0].z = fmaf(sv, d.z, fmaf(c.cii, vz, acc[0].z));
            }
            // the j == i slot (d = 0) left its cii v_i term: taken out again
            const float r2 = 1e-12f, rinv = rsq_fast(r2);
            const float cii = rpy_select(r2, rinv, a, ia).cii;
            acc[0].x = fmaf(-cii, fmaf(Ci[0].x, Ri[0].x, Fi[0].x), acc[0].x);
            acc[0].y = fmaf(-cii, fmaf(Ci[0].y, Ri[0].y, Fi[0].y), acc[0].y);
            acc[0].z = fmaf(-cii, fmaf(Ci[0].z, Ri[0].z, Fi[0].z), acc[0].z);
        }
    } else if (has) {
        const float4 *j0 = &jd[0][l], *j1 = &jd[1][l], *j2 = &jd[2][l];
        // bead h of the lane against j: i-side into acc[h], j-side into g
        auto pair = [&](auto far, int h, const float4 &A, const float4 &B, const float4 &Cq) {
            float3 d = make_float3(A.x - xi[h].x, A.y - xi[h].y, A.z - xi[h].z);
            if (PERIODIC) d = min_image_inv(d, box, ibox);
            // (the clamp is the identity beyond 2a)
            const float r2s = d.x * d.x + d.y * d.y + d.z * d.z;
            const float r2 = decltype(far)::value ? r2s : fmaxf(r2s, 1e-12f);
            const float rinv = rsq_fast(r2);
            const RpyCoef c = decltype(far)::value ? rpy_far(rinv, a) : rpy_select(r2, rinv, a, ia);
            const float cr = c.crr * rinv * rinv;
            // i-side: F_j + beta C_i . R_j
            const float vx = fmaf(Ci[h].x, B.z, A.w), vy = fmaf(Ci[h].y, B.w, B.x), vz = fmaf(Ci[h].z, Cq.x, B.y);
            const float sv = cr * (d.x * vx + d.y * vy + d.z * vz);
            acc[h].x = fmaf(sv, d.x, fmaf(c.cii, vx, acc[h].x));
            acc[h].y = fmaf(sv, d.y, fmaf(c.cii, vy, acc[h].y));
            acc[h].z = fmaf(sv, d.z, fmaf(c.cii, vz, acc[h].z));
            // j-side: F_i + beta C_j . R_i (e e^T is even in e)
            const float wx = fmaf(Cq.y, Ri[h].x, Fi[h].x), wy = fmaf(Cq.z, Ri[h].y, Fi[h].y),
                        wz = fmaf(Cq.w, Ri[h].z, Fi[h].z);
            const float sw = cr * (d.x * wx + d.y * wy + d.z * wz);
            g.x = fmaf(sw, d.x, fmaf(c.cii, wx, g.x));
            g.y = fmaf(sw, d.y, fmaf(c.cii, wy, g.y));
            g.z = fmaf(sw, d.z, fmaf(c.cii, wz, g.z));
        };
        // each step: the lane's beads meet j together, then g moves on
        auto sweep = [&](auto far, auto two) {
            for (int s = 0; s < SYM_CHUNK; s++) {
                const float4 A = j0[s], B = j1[s], Cq = j2[s];
                pair(far, 0, A, B, Cq);
                if (decltype(two)::value) pair(far, 1, A, B, Cq);
                g.x = wave_rol1(g.x);
                g.y = wave_rol1(g.y);
                g.z = wave_rol1(g.z);
            }
        };
        // Open boundaries: when the bounding boxes of the I beads and of
        // chunk J lie more than 2a apart (with a margin over the rounding of
        // r / a), every pair takes the far branch and the selection is
        // skipped (same values).  Slots past the trajectory's end stand in
        // as the chunk's first bead; a one-chunk unit counts its bead twice.
        bool far = false;
        if (!PERIODIC) {
            const float4 xj = jd[0][vj ? l : 0];
            const float3 xb = ni == 2 ? xi[1] : xi[0];
            float lo[6] = {fminf(xi[0].x, xb.x), fminf(xi[0].y, xb.y), fminf(xi[0].z, xb.z), xj.x, xj.y, xj.z};
            float hi[6] = {fmaxf(xi[0].x, xb.x), fmaxf(xi[0].y, xb.y), fmaxf(xi[0].z, xb.z), xj.x, xj.y, xj.z};
#pragma unroll
            for (int m = 1; m < SYM_CHUNK; m <<= 1)
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    lo[k] = fminf(lo[k], __shfl_xor(lo[k], m));
                    hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], m));
                }
            const float gx = fmaxf(fmaxf(lo[3] - hi[0], lo[0] - hi[3]), 0.f);
            const float gy = fmaxf(fmaxf(lo[4] - hi[1], lo[1] - hi[4]), 0.f);
            const float gz = fmaxf(fmaxf(lo[5] - hi[2], lo[2] - hi[5]), 0.f);
            const float lim = 2.002f * a;
            far = __builtin_amdgcn_readfirstlane((int)(gx * gx + gy * gy + gz * gz > lim * lim)) != 0;
        }
        if (ni == 2) {
            if (far) sweep(std::true_type{}, std::true_type{});
            else sweep(std::false_type{}, std::true_type{});
        } else {
            if (far) sweep(std::true_type{}, std::false_type{});
            else sweep(std::false_type{}, std::false_type{});
        }
    }
#pragma unroll
    for (int h = 0; h < 2; h++)
        if (vi[h])
            partial[(size_t)(u.z / SYM_CHUNK) * chunk_stride + bi[h]] =
                ei[h] ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(acc[h].x, acc[h].y, acc[h].z, 0.f);
    if (vj && !diag) {
        // a two-chunk unit's j-side sum covers both chunks: slot I holds
        // it, slot I + 1 zero
        partial[(size_t)(u.y / SYM_CHUNK) * chunk_stride + bj] = ej ? make_float4(0.f, 0.f, 0.f, 0.f)
                                                                     : make_float4(g.x, g.y, g.z, 0.f);
        if (ni == 2) partial[(size_t)(u.y / SYM_CHUNK + 1) * chunk_stride + bj] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Self term, chunk partials in fixed order, and the update of one bead
// (integrateTea_kernel_unlisted, src/bdhitea_kernel.cu:148-213).
__global__ void __launch_bounds__(TEA_TILE)
    k_tea_finish(mte_params P, DevTopo T, DevState Sin, DevState Sout, const float4 *fxyz, const float4 *fang,
                 const float4 *R, const float4 *ci, const float *beta, const int *nchunks, int chunk_stride,
                 const float4 *partial, uint4 *seeds, const float *noise, LatOffsets L) {
    const int b = blockIdx.x * TEA_TILE + threadIdx.x;
    if (b >= T.nb) return;
    const int t = T.tid[b];
    const float bt = beta[t];
    const float4 cc = ci[b], Fi = fxyz[b], Ri = R[b];
    const float w = bt * bt;
    // C_i = 1/sqrt(1 + beta^2 ci) per row (eq. 19)
    const float cx = rsq_fast(1.f + w * cc.x), cy = rsq_fast(1.f + w * cc.y), cz = rsq_fast(1.f + w * cc.z);
    float fx = fmaf(Ri.x, cx, Fi.x), fy = fmaf(Ri.y, cy, Fi.y), fz = fmaf(Ri.z, cz, Fi.z);
    // the bead's state is requested before the partial sums are added
    float4 x = Sin.xyz[b], an = Sin.ang[b];
    const int nc = nchunks[t];
    // partials four at a time: four loads in flight, then added in order
    int c = 0;
    for (; c + 4 <= nc; c += 4) {
        float4 pp[4];
#pragma unroll
        for (int u = 0; u < 4; u++) pp[u] = partial[(size_t)(c + u) * chunk_stride + b];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            fx += pp[u].x;
            fy += pp[u].y;
            fz += pp[u].z;
        }
    }
    for (; c < nc; c++) {
        const float4 pp = partial[(size_t)c * chunk_stride + b];
        fx += pp.x;
        fy += pp.y;
        fz += pp.z;
    }
    float4 ra;
    if (noise) {
        const float *q = noise + 6 * (size_t)b;
        ra = make_float4(q[3], q[4], q[5], 0.f);
    } else {
        uint4 s = seeds[b + T.nb];
        ra = normals4(s);
        seeds[b + T.nb] = s;
    }
    float4 n4 = Sin.fn[b], p1 = Sin.fp1[b], p2 = Sin.fp2[b];
    if (!T.fixed[b] && !T.extra[b]) {
        const float mult = P.dt / P.gammaR;
        x.x = fmaf(mult, fx, x.x);
        x.y = fmaf(mult, fy, x.y);
        x.z = fmaf(mult, fz, x.z);
        const float4 fa = fang[b];
        const float dta = P.dt / (P.gammaTheta * P.alpha), dtt = P.dt / P.gammaTheta;
        const float vta = P.varTheta * sqrtf(P.freeze_temp / P.alpha);
        an.x += dta * fa.x + vta * ra.x;            // fi
        an.z += dta * fa.z + vta * ra.y;            // psi
        an.y += dtt * fa.y + P.varTheta * ra.z;     // theta
        frame_vectors(an.x, an.y, an.z, L, n4, p1, p2);
    }
    Sout.xyz[b] = x;
    Sout.ang[b] = an;
    Sout.fn[b] = n4;
    Sout.fp1[b] = p1;
    Sout.fp2[b] = p2;
}

// ------------------------------------------------------------------ MFMA
// Pair sums on the matrix cores (open boundaries).  With d = r_i - r_j and
// the far-branch coefficients (r > 2a) of D_hat = cii I + c d d^T,
//   cii = p1 w + p3 w^3,  c = p1 w^3 - 3 p3 w^5,  w = 1/r, p1 = 3a/4, p3 = a^3/2,
// the sum over the j of a tile of D_hat_ij (F_j + C'_i . R_j), C' = beta C,
// splits into two scalar kernels times per-j moments about the tile's centre
// Y (s = r_j - Y, b = r_i - Y, d = b - s):
//   sum_j cii (F_j + C'_i . R_j)                          -> moments F, R
//   sum_j c d (d . F_j)  = b (b . S[F]) - b S[s.F] - S[s F^T] b + S[s (s.F)]
//   sum_j c d_al d_ga R_ga = b_al b_ga S[R_ga] - b_al S[s_ga R_ga]
//                            - b_ga S[s_al R_ga] + S[s_al s_ga R_ga]
// with S[m] = sum_j kernel_ij m_j: 6 moment columns on cii and 37 on c.  The
// moments depend on the source tile alone, so k_tea_moments tabulates them
// once per step (split into bf16 hi + lo, a value being hi + lo to 2^-18 of
// itself), and the pair kernel only evaluates the two kernels per pair and
// contracts them with the tables on v_mfma_f32_16x16x32_bf16 (hi hi, hi lo,
// lo hi, accumulated in FP32).  The kernels are symmetric in (i, j), so each
// pair's values serve both beads: tile I's sums over tile J and tile J's
// over tile I, as k_tea_pairs_sym does with its 3 x 3 blocks.
// Tiles are 64 consecutive beads of the spatial (Morton cell) order that
// k_rebuild_order computes when a state is set, a compact patch (on a
// microtubule at most ~24 nm from its centre); any order gives the same
// sums in exact arithmetic, but the rounding of the expansion grows with
// the tile's radius.  Pairs closer than MF_RNEAR (> 2a, so every product
// pair takes the far branch) are summed exactly on the VALU instead: they
// carry the largest kernel values, and beyond MF_RNEAR the terms of the
// expansion exceed the result by at most (1 + |s| / r)^2.  Trajectories
// above REBUILD_ORDER_MAX beads keep the index order, so the engine does
// not offer this form for them.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int MF_TILE = 64;
constexpr int MF_COLS = 64;     // moment columns per bead in the tables: 0-36 on c, 48-53 on cii
constexpr int MF_USED = 43;     // the columns in use, packed (c: 0-36, cii: 37-42)
constexpr int MF_OUT_LD = 68;   // row stride of the LDS result rows (floats)
constexpr float MF_RNEAR = 6.f;

// packed index of moment column c, or -1 for an unused column
__device__ __forceinline__ int mf_packed(int c) { return c < 37 ? c : (c >= 48 && c < 54 ? c - 11 : -1); }

// Moment tables and bead records of one 64-bead tile (one wave; lane = bead
// of the tile): centre Y = mean position of the tile's beads, radius = the
// largest distance of one of them from Y; per bead, with s = r - Y,
//   0-2  s_al (s.F)     3-11 s_al F_be    12 s.F     13-15 F_be
//   16-24 s_al s_ga R_ga  25-33 s_al R_ga  34-36 R_ga             (kernel c)
//   48-50 F   51-53 R                                              (kernel cii)
// Extra beads and slots past the trajectory's end have zero F and R, so all
// their moments vanish.  Tables: mom[tile][hi, lo][column][bead] bf16;
// records rec[tile][position (w = 1 if active), F, R, C'][bead] and
// rbead[tile][bead] (global index, -1 past the end), so that the pair kernel
// reads each tile with coalesced loads that depend on nothing but the unit.
__global__ void __launch_bounds__(MF_TILE)
    k_tea_moments(DevTopo T, DevState S, const float4 *fxyz, const float4 *R, const float4 *ci, const float *beta,
                  const int *order, const int2 *tiles, __bf16 *mom, float4 *ctr, float4 *rec, int *rbead) {
    const int2 tl = tiles[blockIdx.x];   // trajectory, first position
    const int base = T.tbase[tl.x], n = T.tsize[tl.x];
    const int lane = threadIdx.x, p = tl.y + lane;
    const bool in = p < n;
    const int bi = in ? base + order[base + p] : base;
    const float4 x = in ? S.xyz[bi] : make_float4(0.f, 0.f, 0.f, 0.f);
    float sum[4] = {x.x, x.y, x.z, in ? 1.f : 0.f};
#pragma unroll
    for (int m = 1; m < MF_TILE; m <<= 1)
#pragma unroll
        for (int k = 0; k < 4; k++) sum[k] += __shfl_xor(sum[k], m);
    const float cx = sum[0] / sum[3], cy = sum[1] / sum[3], cz = sum[2] / sum[3];
    const bool act = in && !T.extra[bi];
    const float4 F = act ? fxyz[bi] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 Rb = act ? R[bi] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float s[3] = {x.x - cx, x.y - cy, x.z - cz}, Fv[3] = {F.x, F.y, F.z}, Rv[3] = {Rb.x, Rb.y, Rb.z};
    float rad = in ? sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]) : 0.f;
#pragma unroll
    for (int m = 1; m < MF_TILE; m <<= 1) rad = fmaxf(rad, __shfl_xor(rad, m));
    const float sF = s[0] * Fv[0] + s[1] * Fv[1] + s[2] * Fv[2];
    float m[MF_COLS];
#pragma unroll
    for (int c = 0; c < MF_COLS; c++) m[c] = 0.f;
#pragma unroll
    for (int al = 0; al < 3; al++) {
        m[al] = s[al] * sF;
        m[13 + al] = Fv[al];
        m[34 + al] = Rv[al];
        m[48 + al] = Fv[al];
        m[51 + al] = Rv[al];
#pragma unroll
        for (int be = 0; be < 3; be++) {
            m[3 + 3 * al + be] = s[al] * Fv[be];
            m[16 + 3 * al + be] = s[al] * s[be] * Rv[be];
            m[25 + 3 * al + be] = s[al] * Rv[be];
        }
    }
    m[12] = sF;
    __bf16 *hi = mom + (size_t)blockIdx.x * 2 * MF_COLS * MF_TILE, *lo = hi + MF_COLS * MF_TILE;
#pragma unroll
    for (int c = 0; c < MF_COLS; c++) {
        const __bf16 h = (__bf16)m[c];
        hi[c * MF_TILE + lane] = h;
        lo[c * MF_TILE + lane] = (__bf16)(m[c] - (float)h);
    }
    float4 C = make_float4(0.f, 0.f, 0.f, 0.f);
    if (act) {
        const float bt = beta[tl.x], w2b = bt * bt;
        const float4 cc = ci[bi];
        C = make_float4(bt * rsq_fast(1.f + w2b * cc.x), bt * rsq_fast(1.f + w2b * cc.y),
                        bt * rsq_fast(1.f + w2b * cc.z), 0.f);
    }
    float4 *rt = rec + (size_t)blockIdx.x * 4 * MF_TILE;
    rt[lane] = in ? make_float4(x.x, x.y, x.z, act ? 1.f : 0.f) : make_float4(1e6f, 1e6f, 1e6f, 0.f);
    rt[MF_TILE + lane] = F;
    rt[2 * MF_TILE + lane] = Rb;
    rt[3 * MF_TILE + lane] = C;
    rbead[(size_t)blockIdx.x * MF_TILE + lane] = in ? bi : -1;
    if (lane == 0) ctr[blockIdx.x] = make_float4(cx, cy, cz, rad);
}

__device__ __forceinline__ void mf_split8(const float v[8], bf16x8_t &hi, bf16x8_t &lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const __bf16 h = (__bf16)v[q];
        hi[q] = h;
        lo[q] = (__bf16)(v[q] - (float)h);
    }
}

// Sum over one source tile for a target bead from the packed moment sums o
// (MF_USED columns), b = r_target - Y_source and the target's C'.
__device__ __forceinline__ float3 mf_reconstruct(const float *o, const float b[3], const float cp[3]) {
    const float bF = b[0] * o[13] + b[1] * o[14] + b[2] * o[15];
    float f[3];
#pragma unroll
    for (int al = 0; al < 3; al++) {
        // cii (F + C' . R)
        float v = fmaf(cp[al], o[40 + al], o[37 + al]);
        // c d (d . F)
        v += o[al] - (b[0] * o[3 + 3 * al] + b[1] * o[4 + 3 * al] + b[2] * o[5 + 3 * al]) - b[al] * o[12] +
             b[al] * bF;
        // sum_ga C'_ga c d_al d_ga R_ga
#pragma unroll
        for (int ga = 0; ga < 3; ga++)
            v = fmaf(cp[ga],
                     o[16 + 3 * al + ga] - b[ga] * o[25 + 3 * al + ga] - b[al] * o[25 + 4 * ga] +
                         b[al] * b[ga] * o[34 + ga],
                     v);
        f[al] = v;
    }
    return make_float3(f[0], f[1], f[2]);
}

// Pair kernel.  Work unit: (tile I, tile J >= I of the same trajectory,
// partial slots I and J), 256 threads.  Wave w takes the 32 x 32 block
// (i half w / 2, j half w % 2) of the unit's 64 x 64 pairs; lane (g, il)
// evaluates the pairs (i = 16 rb + il, j = 8 g + q) of its block, rb < 2,
// q < 8, which is the A operand layout of two 16 x 32 MFMA tiles (row il,
// k = 8 g + q).  I side: those tiles times tile J's moment tables (16 bytes
// per lane from global memory, requested with the tiles' records before
// anything waits) give the 64 moment sums of each i over the block's 32 j.
// J side: the wave stores its kernel values transposed in LDS (rows j, 8
// consecutive i per 16 bytes, the 16-byte groups XOR-swizzled by row so that
// the four lane groups of a store hit different banks) and reads them back
// as A operands against tile I's tables.  The two halves of each target's
// sums meet in LDS; one thread per target adds them, reconstructs the sum
// over the source tile and writes it to the target's partial slot (I side:
// slot J, J side: slot I), as k_tea_pairs_sym does.  A diagonal unit
// (I == J) sums every ordered pair on the I side alone.  Near pairs
// (r < MF_RNEAR) enter the products as zero and are added exactly per lane,
// on both sides, and reduced across the lanes in fixed order: the result
// does not depend on scheduling.  A unit whose tiles' bounding spheres lie
// more than MF_RNEAR apart holds no near pair and skips those tests.
constexpr int MF_KT_WAVE = 4 * 32 * 32;   // bf16 per wave: kd hi, kd lo, kc hi, kc lo; 32 j x 32 i
struct MfLds {
    float4 rec[2][4][MF_TILE];   // [I, J][position (w: active), F, R, C'][bead]
    int bead[2][MF_TILE];
    float near_i[2][MF_TILE][3];   // exact near-pair sums, I side, per j half
    float near_j[2][MF_TILE][3];   // J side, per i half
    union {
        __bf16 kt[4][MF_KT_WAVE];
        float out[2][2][MF_USED][MF_OUT_LD];   // [side][half][packed column][target]
    } u;
};

__device__ __forceinline__ int mf_kt_index(int type, int jj, int ii) {
    // row jj of 32 i; 16-byte group (8 i) swizzled by (jj / 8) % 4
    return (type * 32 + jj) * 32 + (ii ^ (8 * ((jj >> 3) & 3)));
}

// Kernels of one lane's 16 pairs as MFMA A operands; with NEAR, the pairs
// closer than MF_RNEAR enter as zero and their exact blocks are summed into
// ni (I side, per row block) and nj (J side, per q).
template <bool NEAR>
__device__ __forceinline__ void mf_pairs(const MfLds &L, int ha, int hb, int g, int il, bool diag, float a, float ia,
                                         bf16x8_t (&kA)[2][4], float3 (&ni)[2], float3 (&nj)[8], bool &any_near) {
    const float p1 = 0.75f * a, p3 = 0.5f * a * a * a;
#pragma unroll
    for (int rb = 0; rb < 2; rb++) {
        const int ii = 32 * ha + 16 * rb + il;
        const float4 xi = L.rec[0][0][ii];
        float kd[8], kc[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int jj = 32 * hb + 8 * g + q;
            const float4 xj = L.rec[1][0][jj];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float wv = rsq_fast(NEAR ? fmaxf(r2, 1e-12f) : r2), wv2 = wv * wv, wv3 = wv2 * wv;
            kd[q] = fmaf(p3, wv3, p1 * wv);
            kc[q] = wv3 * fmaf(-3.f * p3, wv2, p1);
            if (NEAR) {
                const bool nearp = r2 < MF_RNEAR * MF_RNEAR;
                kd[q] = nearp ? 0.f : kd[q];
                kc[q] = nearp ? 0.f : kc[q];
                any_near |= nearp;
                if (nearp && xi.w != 0.f && xj.w != 0.f && !(diag && ii == jj)) {
                    // the exact RPY block (both branches), src/bdhitea_kernel.cu:38-58, 104-146
                    const float r2c = fmaxf(r2, 1e-12f), rinv = rsq_fast(r2c);
                    const RpyCoef cf = rpy_select(r2c, rinv, a, ia);
                    const float cr = cf.crr * rinv * rinv;
                    const float4 Ci = L.rec[0][3][ii], Fj = L.rec[1][1][jj], Rj = L.rec[1][2][jj];
                    const float vx = fmaf(Ci.x, Rj.x, Fj.x), vy = fmaf(Ci.y, Rj.y, Fj.y),
                                vz = fmaf(Ci.z, Rj.z, Fj.z);
                    const float sv = cr * (dx * vx + dy * vy + dz * vz);
                    ni[rb].x = fmaf(sv, dx, fmaf(cf.cii, vx, ni[rb].x));
                    ni[rb].y = fmaf(sv, dy, fmaf(cf.cii, vy, ni[rb].y));
                    ni[rb].z = fmaf(sv, dz, fmaf(cf.cii, vz, ni[rb].z));
                    if (!diag) {
                        const float4 Cj = L.rec[1][3][jj], Fi = L.rec[0][1][ii], Ri = L.rec[0][2][ii];
                        const float ux = fmaf(Cj.x, Ri.x, Fi.x), uy = fmaf(Cj.y, Ri.y, Fi.y),
                                    uz = fmaf(Cj.z, Ri.z, Fi.z);
                        const float su = cr * (dx * ux + dy * uy + dz * uz);
                        nj[q].x = fmaf(su, dx, fmaf(cf.cii, ux, nj[q].x));
                        nj[q].y = fmaf(su, dy, fmaf(cf.cii, uy, nj[q].y));
                        nj[q].z = fmaf(su, dz, fmaf(cf.cii, uz, nj[q].z));
                    }
                }
            }
        }
        mf_split8(kd, kA[rb][0], kA[rb][1]);
        mf_split8(kc, kA[rb][2], kA[rb][3]);
    }
}

// kernel x table product of one 16-column tile: hi hi + hi lo + lo hi
__device__ __forceinline__ f32x4_t mf_product(const bf16x8_t &khi, const bf16x8_t &klo, const bf16x8_t &bh,
                                              const bf16x8_t &bl) {
    const f32x4_t zero4 = {0.f, 0.f, 0.f, 0.f};
    f32x4_t s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(khi, bh, zero4, 0, 0, 0);
    s = __builtin_amdgcn_mfma_f32_16x16x32_bf16(khi, bl, s, 0, 0, 0);
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(klo, bh, s, 0, 0, 0);
}

__global__ void __launch_bounds__(256)
    k_tea_pairs_mfma(mte_params P, const int4 *units, const __bf16 *mom, const float4 *ctr, const float4 *rec,
                     const int *rbead, int chunk_stride, float4 *partial) {
    __shared__ MfLds L;
    const int4 u = units[blockIdx.x];   // tile I, tile J, slot I, slot J
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, il = lane & 15;
    const int ha = w >> 1, hb = w & 1;   // i half, j half of this wave's block
    const bool diag = u.x == u.y;
    // requests that depend on the unit alone: both tiles' records, this
    // lane's fragments of both tables (16 bytes each), the two centres
    float4 rv[4];
    int rbv = -1;
    const int side0 = tid >> 6, k0 = tid & 63;
    if (tid < 2 * MF_TILE) {
        const float4 *rt = rec + (size_t)(side0 ? u.y : u.x) * 4 * MF_TILE;
#pragma unroll
        for (int c = 0; c < 4; c++) rv[c] = rt[c * MF_TILE + k0];
        rbv = rbead[(size_t)(side0 ? u.y : u.x) * MF_TILE + k0];
    }
    bf16x8_t tb[2][4][2];   // [I side: tile J's table, J side: tile I's table][column tile][hi, lo]
#pragma unroll
    for (int sd = 0; sd < 2; sd++) {
        if (sd == 1 && diag) break;
        const __bf16 *mt = mom + (size_t)(sd ? u.x : u.y) * 2 * MF_COLS * MF_TILE;
        const int kk = 32 * (sd ? ha : hb) + 8 * g;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int col = 16 * t + il;
            tb[sd][t][0] = *reinterpret_cast<const bf16x8_t *>(mt + col * MF_TILE + kk);
            tb[sd][t][1] = *reinterpret_cast<const bf16x8_t *>(mt + (MF_COLS + col) * MF_TILE + kk);
        }
    }
    const float4 YI = ctr[u.x], YJ = ctr[u.y];
    if (tid < 2 * MF_TILE) {
#pragma unroll
        for (int c = 0; c < 4; c++) L.rec[side0][c][k0] = rv[c];
        L.bead[side0][k0] = rbv;
    }
    __syncthreads();
    const float a = P.tea_a, ia = rcp_fast(P.tea_a);
    bf16x8_t kA[2][4];   // [rb][kd hi, kd lo, kc hi, kc lo]
    float3 ni[2] = {make_float3(0.f, 0.f, 0.f), make_float3(0.f, 0.f, 0.f)};
    float3 nj[8];
#pragma unroll
    for (int q = 0; q < 8; q++) nj[q] = make_float3(0.f, 0.f, 0.f);
    bool any_near = false;
    const float gap = sqrtf((YI.x - YJ.x) * (YI.x - YJ.x) + (YI.y - YJ.y) * (YI.y - YJ.y) +
                            (YI.z - YJ.z) * (YI.z - YJ.z)) - YI.w - YJ.w;
    const bool near_unit = diag || !(gap > MF_RNEAR + 0.01f);
    if (near_unit) mf_pairs<true>(L, ha, hb, g, il, diag, a, ia, kA, ni, nj, any_near);
    else mf_pairs<false>(L, ha, hb, g, il, diag, a, ia, kA, ni, nj, any_near);
    // I side: the block's pairs times tile J's tables (k = j of half hb)
    f32x4_t acc[2][4], accj[2][4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int kh = t < 3 ? 2 : 0;   // kernel c for the 48 columns 0-47, cii for 48-63
#pragma unroll
        for (int rb = 0; rb < 2; rb++) acc[rb][t] = mf_product(kA[rb][kh], kA[rb][kh + 1], tb[0][t][0], tb[0][t][1]);
    }
    // J side, first the transposed kernels of this wave's block
    if (!diag) {
        __bf16 *kt = L.u.kt[w];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int q = 0; q < 8; q++)
#pragma unroll
                for (int ty = 0; ty < 4; ty++) kt[mf_kt_index(ty, 8 * g + q, 16 * rb + il)] = kA[rb][ty][q];
    }
    __syncthreads();
    if (!diag) {
        const __bf16 *kt = L.u.kt[w];
        bf16x8_t kB[2][4];
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int ty = 0; ty < 4; ty++)
                kB[rb][ty] = *reinterpret_cast<const bf16x8_t *>(kt + mf_kt_index(ty, 16 * rb + il, 8 * g));
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int kh = t < 3 ? 2 : 0;
#pragma unroll
            for (int rb = 0; rb < 2; rb++)
                accj[rb][t] = mf_product(kB[rb][kh], kB[rb][kh + 1], tb[1][t][0], tb[1][t][1]);
        }
    }
    // near sums: I side over the four lane groups, J side over the 16 lanes of a row
    if (near_unit && __ballot(any_near) != 0) {
#pragma unroll
        for (int rb = 0; rb < 2; rb++)
#pragma unroll
            for (int m = 16; m < 64; m <<= 1) {
                ni[rb].x += __shfl_xor(ni[rb].x, m);
                ni[rb].y += __shfl_xor(ni[rb].y, m);
                ni[rb].z += __shfl_xor(ni[rb].z, m);
            }
        if (!diag)
#pragma unroll
            for (int q = 0; q < 8; q++)
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) {
                    nj[q].x += __shfl_xor(nj[q].x, m);
                    nj[q].y += __shfl_xor(nj[q].y, m);
                    nj[q].z += __shfl_xor(nj[q].z, m);
                }
    }
    __syncthreads();   // every wave is done with its kt before out overwrites them
    if (g == 0)
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            float *o = L.near_i[hb][32 * ha + 16 * rb + il];
            o[0] = ni[rb].x;
            o[1] = ni[rb].y;
            o[2] = ni[rb].z;
        }
    if (il == 0 && !diag)
#pragma unroll
        for (int q = 0; q < 8; q++) {
            float *o = L.near_j[ha][32 * hb + 8 * g + q];
            o[0] = nj[q].x;
            o[1] = nj[q].y;
            o[2] = nj[q].z;
        }
    // lane (g, il) holds rows 4 g + r of column il of each 16-column tile
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int pc = mf_packed(16 * t + il);
        if (pc < 0) continue;
#pragma unroll
        for (int rb = 0; rb < 2; rb++) {
            *reinterpret_cast<f32x4_t *>(&L.u.out[0][hb][pc][32 * ha + 16 * rb + 4 * g]) = acc[rb][t];
            if (!diag) *reinterpret_cast<f32x4_t *>(&L.u.out[1][ha][pc][32 * hb + 16 * rb + 4 * g]) = accj[rb][t];
        }
    }
    __syncthreads();
    if (tid >= 2 * MF_TILE) return;
    const int side = tid >> 6, k = tid & 63;
    if (side == 1 && diag) return;
    const int bi = L.bead[side][k];
    if (bi < 0) return;
    float3 f = make_float3(0.f, 0.f, 0.f);
    const float4 xt = L.rec[side][0][k];
    if (xt.w != 0.f) {
        float o[MF_USED];
#pragma unroll
        for (int c = 0; c < MF_USED; c++) o[c] = L.u.out[side][0][c][k] + L.u.out[side][1][c][k];
        const float4 Y = side ? YI : YJ;
        const float4 cc = L.rec[side][3][k];
        const float b[3] = {xt.x - Y.x, xt.y - Y.y, xt.z - Y.z}, cp[3] = {cc.x, cc.y, cc.z};
        f = mf_reconstruct(o, b, cp);
        const float *n0 = side ? L.near_j[0][k] : L.near_i[0][k], *n1 = side ? L.near_j[1][k] : L.near_i[1][k];
        f.x += n0[0] + n1[0];
        f.y += n0[1] + n1[1];
        f.z += n0[2] + n1[2];
    }
    const int slot = side ? u.z : u.w;
    partial[(size_t)slot * chunk_stride + bi] = make_float4(f.x, f.y, f.z, 0.f);
}

hipError_t launch_tea_eps(mte_engine *e) {
    if (e->tune.tea_eps == MTE_TEA_EPS_ONEPASS) {
        auto kern = e->periodic ? (e->any_extra ? k_tea_eps_sym<true, true> : k_tea_eps_sym<true, false>)
                                : (e->any_extra ? k_tea_eps_sym<false, true> : k_tea_eps_sym<false, false>);
        hipLaunchKernelGGL(kern, dim3(e->tea.eps_nunits), dim3(64 * EPS_WAVES), 0, e->stream, e->P, e->T, e->S[e->cur],
                           e->tea.eps_units, e->tea.eps_ngrp, e->tea.eps_part);
        hipLaunchKernelGGL(k_tea_eps_reduce, dim3((e->nb + 255) / 256), dim3(256), 0, e->stream, e->T,
                           e->tea.eps_ngrp, e->tea.eps_part, e->tea.ci, e->tea.eps);
    } else if (e->periodic)
        hipLaunchKernelGGL(k_tea_eps<true>, dim3(e->ntiles), dim3(TEA_TILE), 0, e->stream, e->P, e->T,
                           e->S[e->cur], e->tiles, e->tea.ci, e->tea.eps);
    else
        hipLaunchKernelGGL(k_tea_eps<false>, dim3(e->ntiles), dim3(TEA_TILE), 0, e->stream, e->P, e->T,
                           e->S[e->cur], e->tiles, e->tea.ci, e->tea.eps);
    hipLaunchKernelGGL(k_tea_beta, dim3(e->ntraj), dim3(256), 0, e->stream, e->P, e->T, e->tea.eps, e->tea.beta,
                       e->tea.stats, e->tea.status);
    return hipGetLastError();
}

template <int CHUNK>
static void launch_pairs_chunk(mte_engine *e) {
    if (e->periodic)
        hipLaunchKernelGGL((k_tea_pairs<CHUNK, true>), dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P,
                           e->T, e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb,
                           e->tea.partial);
    else
        hipLaunchKernelGGL((k_tea_pairs<CHUNK, false>), dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P,
                           e->T, e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb,
                           e->tea.partial);
}

hipError_t launch_tea_pairs(mte_engine *e) {
    // the kernel's chunk must be the one the work units were built for
    if (e->tea.unit_chunk == MTE_TEA_SYMMETRIC) {
        const int nwg = e->tea.nunits;
        if (e->periodic)
            hipLaunchKernelGGL(k_tea_pairs_sym<true>, dim3(nwg), dim3(SYM_CHUNK), 0, e->stream, e->P, e->T,
                               e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units,
                               e->tea.nunits, e->nb, e->tea.partial);
        else
            hipLaunchKernelGGL(k_tea_pairs_sym<false>, dim3(nwg), dim3(SYM_CHUNK), 0, e->stream, e->P, e->T,
                               e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units,
                               e->tea.nunits, e->nb, e->tea.partial);
        return hipGetLastError();
    }
    if (e->tea.unit_chunk == MTE_TEA_MFMA) {
        if (e->periodic) return hipErrorInvalidValue;   // minimum images are not in the moment form
        hipLaunchKernelGGL(k_tea_moments, dim3(e->tea.ntiles), dim3(MF_TILE), 0, e->stream, e->T, e->S[e->cur],
                           e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->rb_order, e->tea.tiles,
                           static_cast<__bf16 *>(e->tea.mom), e->tea.ctr, e->tea.rec, e->tea.rbead);
        hipLaunchKernelGGL(k_tea_pairs_mfma, dim3(e->tea.nunits), dim3(256), 0, e->stream, e->P, e->tea.units,
                           static_cast<const __bf16 *>(e->tea.mom), e->tea.ctr, e->tea.rec, e->tea.rbead, e->nb,
                           e->tea.partial);
        return hipGetLastError();
    }
    switch (e->tea.unit_chunk) {
    case 64: launch_pairs_chunk<64>(e); break;
    case 128: launch_pairs_chunk<128>(e); break;
    case 256: launch_pairs_chunk<256>(e); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tea_finish(mte_engine *e, const float *noise_step) {
    hipLaunchKernelGGL(k_tea_finish, dim3((e->nb + TEA_TILE - 1) / TEA_TILE), dim3(TEA_TILE), 0, e->stream, e->P,
                       e->T, e->S[e->cur], e->S[1 - e->cur], e->fxyz, e->fang, e->tea.R, e->tea.ci, e->tea.beta,
                       e->tea.nchunks, e->nb, e->tea.partial, e->seeds, noise_step, e->L);
    e->cur = 1 - e->cur;
    return hipGetLastError();
}

}  // namespace mte
