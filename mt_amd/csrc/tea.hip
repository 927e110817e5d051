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
// (trajectory, 1024-bead group I, group J >= I), one workgroup of 16 waves;
// wave w takes chunk c_I = 16 I + w (64 beads) against each 64-bead chunk c_J
// of group J in turn, staged in LDS:
//   c_J > c_I  both sides, lane l meeting j = (l + s) mod 64 at step s, the
//              j-side sums carried along by a one-lane rotation (wave_rol1),
//   c_J = c_I  one-sided over the chunk, the j == i term taken out after,
//   c_J < c_I  nothing (that pair of chunks is the wave c_J's).
// The waves' j-side sums of a chunk are added in wave order in LDS.  Output
// (deterministic, no atomics): i-side sums of a bead of group H into
// part[0][J][.] for J >= H, j-side sums into part[1][I][.] for I <= H; the
// pair weight is 1 for two non-extra beads and 0 otherwise.
constexpr int EPS_CHUNKS = EPS_GROUP / 64;
template <bool PERIODIC>
__global__ void __launch_bounds__(EPS_GROUP)
    k_tea_eps_sym(mte_params P, DevTopo T, DevState S, const int4 *units, int ngrp, float *part) {
    __shared__ float4 jx[2 * 64];               // chunk c_J: x, weight; twice, no wrap
    __shared__ float jred[EPS_CHUNKS][5][64];   // j-side sums of each wave
    const int4 u = units[blockIdx.x];
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int base = T.tbase[u.x], n = T.tsize[u.x], nb = T.nb;
    const int nch = (n + 63) / 64;
    const float box = T.tbox[u.x], ibox = inv_side(box);
    const float a = P.tea_a, ia = rcp_fast(P.tea_a);
    const int cI = u.y * EPS_CHUNKS + w;
    const int li = cI * 64 + l;
    const bool vi = li < n;
    float3 xi = make_float3(0.f, 0.f, 0.f);
    float wi = 0.f;
    if (vi) {
        xi = f3(S.xyz[base + li]);
        wi = T.extra[base + li] ? 0.f : 1.f;
    }
    float ai[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    const int c0 = u.z * EPS_CHUNKS, c1 = min(c0 + EPS_CHUNKS, nch);
    for (int cJ = c0; cJ < c1; cJ++) {
        if (threadIdx.x < 128) {
            const int lj = cJ * 64 + (threadIdx.x % 64);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lj < n) {
                v = S.xyz[base + lj];
                v.w = T.extra[base + lj] ? 0.f : 1.f;
            }
            jx[threadIdx.x] = v;
        }
        __syncthreads();
        float gj[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
        // one pair's five terms into the sums t (and t2 when both sides); a
        // wave with no pair inside 2a (margin over the rounding of r / a)
        // takes the far branch alone, which is the branch rpy_select picks
        const float near2 = (2.002f * a) * (2.002f * a);
        auto pair = [&](float4 xj, float *t, float *t2) {
            float3 d = make_float3(xj.x - xi.x, xj.y - xi.y, xj.z - xi.z);
            if (PERIODIC) d = min_image_inv(d, box, ibox);
            const float r2 = fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, 1e-12f);
            const float rinv = rsq_fast(r2);
            RpyCoef c;
            if (__any(r2 <= near2)) c = rpy_select(r2, rinv, a, ia);
            else c = rpy_far(rinv, a);
            const float ww = wi * xj.w;
            const float ex = d.x * rinv, ey = d.y * rinv, ez = d.z * rinv;
            const float cw = ww * c.cii;
            const float k = ww * c.crr * fmaf(2.f, c.cii, c.crr);
            const float se = ex + ey + ez;
            const float tall = fmaf(ww * c.crr, se * se, 3.f * cw);
            const float qx = ex * ex, qy = ey * ey, qz = ez * ez;
            t[0] = fmaf(cw, c.cii, t[0]);
            t[1] = fmaf(k, qx, t[1]);
            t[2] = fmaf(k, qy, t[2]);
            t[3] = fmaf(k, qz, t[3]);
            t[4] += tall;
            if (t2) {
                t2[0] = fmaf(cw, c.cii, t2[0]);
                t2[1] = fmaf(k, qx, t2[1]);
                t2[2] = fmaf(k, qy, t2[2]);
                t2[3] = fmaf(k, qz, t2[3]);
                t2[4] += tall;
            }
        };
        if (cI < nch && cJ > cI) {
            for (int s = 0; s < 64; s++) {
                pair(jx[l + s], ai, gj);
#pragma unroll
                for (int k = 0; k < 5; k++) gj[k] = wave_rol1(gj[k]);
            }
        } else if (cI < nch && cJ == cI) {
            const int jn = min(64, n - cJ * 64);
            for (int q = 0; q < jn; q++) pair(jx[q], ai, nullptr);
            if (wi != 0.f) {
                // the j == i slot entered as a coincident bead (d = 0):
                // cii^2 into the first sum and 3 cii into the last, taken out
                const float r2 = 1e-12f, rinv = rsq_fast(r2);
                const float cii = rpy_select(r2, rinv, a, ia).cii;
                ai[0] = fmaf(-cii, cii, ai[0]);
                ai[4] = fmaf(-3.f, cii, ai[4]);
            }
        }
#pragma unroll
        for (int k = 0; k < 5; k++) jred[w][k][l] = gj[k];
        __syncthreads();
        if (threadIdx.x < 5 * 64) {
            const int k = threadIdx.x / 64, q = threadIdx.x % 64;
            const int lj = cJ * 64 + q;
            if (lj < n) {
                float sum = 0.f;
                for (int v = 0; v < EPS_CHUNKS; v++) sum += jred[v][k][q];
                part[(((size_t)1 * ngrp + u.y) * 5 + k) * nb + base + lj] = sum;
            }
        }
        __syncthreads();
    }
    if (vi) {
#pragma unroll
        for (int k = 0; k < 5; k++) part[(((size_t)0 * ngrp + u.z) * 5 + k) * nb + base + li] = ai[k];
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
    float s[5];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        float v = 0.f;
        for (int j = h; j < g; j++) v += part[(((size_t)0 * ngrp + j) * 5 + k) * nb + b];
        for (int i = 0; i <= h; i++) v += part[(((size_t)1 * ngrp + i) * 5 + k) * nb + b];
        s[k] = v;
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
// it.  Off the diagonal, lane l takes j = (l + s) mod 64 at step s: every
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
    // bead of this lane in chunk I ...
    const int li = u.y + l, bi = base + li;
    const bool vi = has && li < n;
    const bool ei = vi && T.extra[bi];
    float3 xi = make_float3(0.f, 0.f, 0.f), Fi = xi, Ri = xi, Ci = xi;
    if (vi) {
        xi = f3(S.xyz[bi]);
        if (!ei) {
            Fi = f3(fxyz[bi]);
            Ri = f3(R[bi]);
            const float4 cc = ci[bi];
            Ci = make_float3(bt * rsq_fast(1.f + w2 * cc.x), bt * rsq_fast(1.f + w2 * cc.y),
                             bt * rsq_fast(1.f + w2 * cc.z));
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
    float3 acc = make_float3(0.f, 0.f, 0.f), g = acc;
    const bool diag = u.y == u.z;
    if (has && diag) {
        if (vi && !ei) {
            const int jn = min(SYM_CHUNK, n - u.z);
            for (int q = 0; q < jn; q++) {
                const float4 A = jd[0][q], B = jd[1][q], Cq = jd[2][q];
                float3 d = make_float3(A.x - xi.x, A.y - xi.y, A.z - xi.z);
                if (PERIODIC) d = min_image_inv(d, box, ibox);
                const float r2 = fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, 1e-12f);
                const float rinv = rsq_fast(r2);
                const RpyCoef c = rpy_select(r2, rinv, a, ia);
                const float vx = fmaf(Ci.x, B.z, A.w), vy = fmaf(Ci.y, B.w, B.x), vz = fmaf(Ci.z, Cq.x, B.y);
                const float sv = (c.crr * rinv * rinv) * (d.x * vx + d.y * vy + d.z * vz);
                acc.x = fmaf(sv, d.x, fmaf(c.cii, vx, acc.x));
                acc.y = fmaf(sv, d.y, fmaf(c.cii, vy, acc.y));
                acc.z = fmaf(sv, d.z, fmaf(c.cii, vz, acc.z));
            }
            // the j == i slot (d = 0) left its cii v_i term: taken out again
            const float r2 = 1e-12f, rinv = rsq_fast(r2);
            const float cii = rpy_select(r2, rinv, a, ia).cii;
            acc.x = fmaf(-cii, fmaf(Ci.x, Ri.x, Fi.x), acc.x);
            acc.y = fmaf(-cii, fmaf(Ci.y, Ri.y, Fi.y), acc.y);
            acc.z = fmaf(-cii, fmaf(Ci.z, Ri.z, Fi.z), acc.z);
        }
    } else if (has) {
        const float4 *j0 = &jd[0][l], *j1 = &jd[1][l], *j2 = &jd[2][l];
        auto sweep = [&](auto far) {
            for (int s = 0; s < SYM_CHUNK; s++) {
                const float4 A = j0[s], B = j1[s], Cq = j2[s];
                float3 d = make_float3(A.x - xi.x, A.y - xi.y, A.z - xi.z);
                if (PERIODIC) d = min_image_inv(d, box, ibox);
                // (the clamp is the identity beyond 2a)
                const float r2s = d.x * d.x + d.y * d.y + d.z * d.z;
                const float r2 = decltype(far)::value ? r2s : fmaxf(r2s, 1e-12f);
                const float rinv = rsq_fast(r2);
                const RpyCoef c = decltype(far)::value ? rpy_far(rinv, a) : rpy_select(r2, rinv, a, ia);
                const float cr = c.crr * rinv * rinv;
                // i-side: F_j + beta C_i . R_j
                const float vx = fmaf(Ci.x, B.z, A.w), vy = fmaf(Ci.y, B.w, B.x), vz = fmaf(Ci.z, Cq.x, B.y);
                const float sv = cr * (d.x * vx + d.y * vy + d.z * vz);
                acc.x = fmaf(sv, d.x, fmaf(c.cii, vx, acc.x));
                acc.y = fmaf(sv, d.y, fmaf(c.cii, vy, acc.y));
                acc.z = fmaf(sv, d.z, fmaf(c.cii, vz, acc.z));
                // j-side: F_i + beta C_j . R_i (e e^T is even in e)
                const float wx = fmaf(Cq.y, Ri.x, Fi.x), wy = fmaf(Cq.z, Ri.y, Fi.y), wz = fmaf(Cq.w, Ri.z, Fi.z);
                const float sw = cr * (d.x * wx + d.y * wy + d.z * wz);
                g.x = wave_rol1(fmaf(sw, d.x, fmaf(c.cii, wx, g.x)));
                g.y = wave_rol1(fmaf(sw, d.y, fmaf(c.cii, wy, g.y)));
                g.z = wave_rol1(fmaf(sw, d.z, fmaf(c.cii, wz, g.z)));
            }
        };
        // Open boundaries: when the bounding boxes of the two chunks lie more
        // than 2a apart (with a margin over the rounding of r / a), every
        // pair takes the far branch and the selection is skipped (same
        // values).  Slots past the trajectory's end stand in as the chunk's
        // first bead.
        bool far = false;
        if (!PERIODIC) {
            const float4 xj = jd[0][vj ? l : 0];
            float lo[6] = {xi.x, xi.y, xi.z, xj.x, xj.y, xj.z}, hi[6] = {xi.x, xi.y, xi.z, xj.x, xj.y, xj.z};
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
        if (far) sweep(std::true_type{});
        else sweep(std::false_type{});
    }
    if (vi) partial[(size_t)(u.z / SYM_CHUNK) * chunk_stride + bi] = ei ? make_float4(0.f, 0.f, 0.f, 0.f)
                                                                         : make_float4(acc.x, acc.y, acc.z, 0.f);
    if (vj && !diag)
        partial[(size_t)(u.y / SYM_CHUNK) * chunk_stride + bj] = ej ? make_float4(0.f, 0.f, 0.f, 0.f)
                                                                     : make_float4(g.x, g.y, g.z, 0.f);
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
// Pair sums on the matrix cores (open boundaries).  With d = r_j - r_i,
// c(r) = crr / r^2 and the far-branch coefficients (r > 2a)
//   cii = p1 w + p3 w^3,  c = p1 w^3 - 3 p3 w^5,  w = 1/r, p1 = 3a/4, p3 = a^3/2,
// the sum over j of D_hat_ij (F_j + C'_i . R_j), C' = beta C, splits into
// scalar kernels times per-j moments taken about a centre X (r = r_j - X,
// a = r_i - X):
//   sum_j cii (F_j + C'_i . R_j)                        -> moments F, R
//   sum_j c d (d . F_j)  = S[r (r.F)] - S[r F^T] a - a S[r.F] + a (a . S[F])
//   sum_j c d_al d_ga R_ga = S[r_al r_ga R_ga] - a_ga S[r_al R_ga]
//                            - a_al S[r_ga R_ga] + a_al a_ga S[R_ga]
// with S[m] = sum_j kernel_ij m_j: 6 moment columns on cii and 37 on c, i.e.
// four 16-column products per 16 x 32 block of pairs, each as three
// v_mfma_f32_16x16x32_bf16 on split operands (hi hi, hi lo, lo hi; a value
// is hi + lo to 2^-18 of itself), accumulated in FP32.
// Work unit: (trajectory, 64-bead I tile, 64-bead J chunk), four waves of
// 16 i each; X = the mean position of the I tile.  The I tiles take the
// beads in the spatial (Morton cell) order k_rebuild_order computes when a
// state is set, so a tile is a compact patch (at most ~18 nm from its centre
// on a microtubule) rather than 64 beads along a protofilament; any order
// gives the same sums in exact arithmetic, but the rounding of the moments
// grows with the tile's radius.  Pairs closer than MF_RNEAR (which is > 2a,
// so every product pair takes the far branch) are summed exactly on the VALU
// instead: they carry the largest kernel values, and beyond MF_RNEAR the
// terms of the expansion exceed the result by at most (1 + |a| / r)^2 for a
// bead at a from the centre.  Trajectories above REBUILD_ORDER_MAX beads keep
// the index order, so the engine does not offer this kernel for them.
// j == i, extra beads and slots past the trajectory end enter as zero.
// Partial sums of chunk J go to slot J of bead i, as k_tea_pairs<64> writes
// them.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
constexpr int MF_TILE = 64;
constexpr int MF_COLS = 64;          // moment columns: 0-47 on c (37 used), 48-63 on cii (6 used)
constexpr float MF_RNEAR = 16.f;

__device__ __forceinline__ void mf_split8(const float v[8], bf16x8_t &hi, bf16x8_t &lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const __bf16 h = (__bf16)v[q];
        hi[q] = h;
        lo[q] = (__bf16)(v[q] - (float)h);
    }
}

__global__ void __launch_bounds__(256)
    k_tea_pairs_mfma(mte_params P, DevTopo T, DevState S, const float4 *fxyz, const float4 *R, const float4 *ci,
                     const float *beta, const int4 *units, const int *order, int chunk_stride, float4 *partial) {
    __shared__ float4 jx[MF_TILE], jf[MF_TILE], jr[MF_TILE];   // w of jx: 1 for a bead that takes part
    __shared__ __attribute__((aligned(16))) __bf16 mhi[MF_COLS][MF_TILE];
    __shared__ __attribute__((aligned(16))) __bf16 mlo[MF_COLS][MF_TILE];
    __shared__ float out[4][16][MF_COLS + 1];
    __shared__ float nr[4][4][16][3];
    __shared__ float4 xsum[4];
    const int4 u = units[blockIdx.x];   // traj, i0, j0, chunk index
    const int base = T.tbase[u.x], n = T.tsize[u.x];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, il = lane & 15;
    const float bt = beta[u.x], w2b = bt * bt;
    // the I tile's centre
    {
        float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
        if (tid < MF_TILE && u.y + tid < n) {
            x = S.xyz[base + order[base + u.y + tid]];
            x.w = 1.f;
        }
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            x.x += __shfl_xor(x.x, m);
            x.y += __shfl_xor(x.y, m);
            x.z += __shfl_xor(x.z, m);
            x.w += __shfl_xor(x.w, m);
        }
        if (tid == 0) xsum[0] = x;
    }
    // the J chunk
    if (tid < MF_TILE) {
        const int lj = u.z + tid, bj = base + lj;
        float4 x = make_float4(0.f, 0.f, 0.f, 0.f), f = x, r = x;
        if (lj < n) {
            x = S.xyz[bj];
            x.w = T.extra[bj] ? 0.f : 1.f;
            if (x.w != 0.f) {
                f = fxyz[bj];
                r = R[bj];
            }
        }
        jx[tid] = x;
        jf[tid] = f;
        jr[tid] = r;
    }
    __syncthreads();
    const float4 xs = xsum[0];
    const float cx0 = xs.x / xs.w, cy0 = xs.y / xs.w, cz0 = xs.z / xs.w;
    // moments about the centre: thread (column group tid / 64, j = tid % 64)
    {
        const int j = lane, grp = w;
        const float4 xj = jx[j], F = jf[j], Rj = jr[j];
        const float r0 = xj.x - cx0, r1 = xj.y - cy0, r2 = xj.z - cz0;
        const float rr[3] = {r0, r1, r2}, Fv[3] = {F.x, F.y, F.z}, Rv[3] = {Rj.x, Rj.y, Rj.z};
        float m[16];
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = 0.f;
        if (xj.w != 0.f) {
            const float rF = r0 * Fv[0] + r1 * Fv[1] + r2 * Fv[2];
            if (grp == 0) {
#pragma unroll
                for (int al = 0; al < 3; al++) m[al] = rr[al] * rF;
#pragma unroll
                for (int al = 0; al < 3; al++)
#pragma unroll
                    for (int be = 0; be < 3; be++) m[3 + 3 * al + be] = rr[al] * Fv[be];
                m[12] = rF;
#pragma unroll
                for (int be = 0; be < 3; be++) m[13 + be] = Fv[be];
            } else if (grp == 1) {
#pragma unroll
                for (int al = 0; al < 3; al++)
#pragma unroll
                    for (int ga = 0; ga < 3; ga++) m[3 * al + ga] = rr[al] * rr[ga] * Rv[ga];
#pragma unroll
                for (int k = 0; k < 7; k++) m[9 + k] = rr[k / 3] * Rv[k % 3];
            } else if (grp == 2) {
                m[0] = rr[2] * Rv[1];
                m[1] = rr[2] * Rv[2];
#pragma unroll
                for (int ga = 0; ga < 3; ga++) m[2 + ga] = Rv[ga];
            } else {
#pragma unroll
                for (int be = 0; be < 3; be++) {
                    m[be] = Fv[be];
                    m[3 + be] = Rv[be];
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const __bf16 h = (__bf16)m[q];
            mhi[16 * grp + q][j] = h;
            mlo[16 * grp + q][j] = (__bf16)(m[q] - (float)h);
        }
    }
    __syncthreads();
    // this lane's bead: row il of the wave's 16
    const int pi = u.y + 16 * w + il;
    const int li = pi < n ? order[base + pi] : n, bi = base + li;
    const bool vi = li < n && !T.extra[bi];
    float3 xi = make_float3(cx0, cy0, cz0), Cp = make_float3(0.f, 0.f, 0.f);
    if (vi) {
        xi = f3(S.xyz[bi]);
        const float4 cc = ci[bi];
        Cp = make_float3(bt * rsq_fast(1.f + w2b * cc.x), bt * rsq_fast(1.f + w2b * cc.y),
                         bt * rsq_fast(1.f + w2b * cc.z));
    }
    const float a = P.tea_a, ia = rcp_fast(P.tea_a);
    const float p1 = 0.75f * a, p3 = 0.5f * a * a * a;
    f32x4_t accA = {0.f, 0.f, 0.f, 0.f}, accC[3];
#pragma unroll
    for (int t = 0; t < 3; t++) accC[t] = accA;
    float3 nacc = make_float3(0.f, 0.f, 0.f);
#pragma unroll
    for (int kb = 0; kb < 2; kb++) {
        float kc[8], kd[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int jj = kb * 32 + 8 * g + q;
            const float4 xj = jx[jj];
            const float dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const bool nearp = r2 < MF_RNEAR * MF_RNEAR;
            const float wv = rsq_fast(fmaxf(r2, 1e-12f)), wv2 = wv * wv, wv3 = wv2 * wv;
            kd[q] = nearp ? 0.f : fmaf(p3, wv3, p1 * wv);
            kc[q] = nearp ? 0.f : fmaf(-3.f * p3, wv3 * wv2, p1 * wv3);
            if (nearp && vi && xj.w != 0.f && u.z + jj != li) {
                // the exact RPY block (both branches), src/bdhitea_kernel.cu:38-58, 104-146
                const float r2c = fmaxf(r2, 1e-12f), rinv = rsq_fast(r2c);
                const RpyCoef c = rpy_select(r2c, rinv, a, ia);
                const float4 F = jf[jj], Rj = jr[jj];
                const float vx = fmaf(Cp.x, Rj.x, F.x), vy = fmaf(Cp.y, Rj.y, F.y), vz = fmaf(Cp.z, Rj.z, F.z);
                const float sv = (c.crr * rinv * rinv) * (dx * vx + dy * vy + dz * vz);
                nacc.x = fmaf(sv, dx, fmaf(c.cii, vx, nacc.x));
                nacc.y = fmaf(sv, dy, fmaf(c.cii, vy, nacc.y));
                nacc.z = fmaf(sv, dz, fmaf(c.cii, vz, nacc.z));
            }
        }
        bf16x8_t dhi, dlo, chi, clo;
        mf_split8(kd, dhi, dlo);
        mf_split8(kc, chi, clo);
        const int k0 = kb * 32 + 8 * g;
        {
            const bf16x8_t bh = *reinterpret_cast<const bf16x8_t *>(&mhi[48 + il][k0]);
            const bf16x8_t bl = *reinterpret_cast<const bf16x8_t *>(&mlo[48 + il][k0]);
            accA = __builtin_amdgcn_mfma_f32_16x16x32_bf16(dhi, bh, accA, 0, 0, 0);
            accA = __builtin_amdgcn_mfma_f32_16x16x32_bf16(dhi, bl, accA, 0, 0, 0);
            accA = __builtin_amdgcn_mfma_f32_16x16x32_bf16(dlo, bh, accA, 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const bf16x8_t bh = *reinterpret_cast<const bf16x8_t *>(&mhi[16 * t + il][k0]);
            const bf16x8_t bl = *reinterpret_cast<const bf16x8_t *>(&mlo[16 * t + il][k0]);
            accC[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(chi, bh, accC[t], 0, 0, 0);
            accC[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(chi, bl, accC[t], 0, 0, 0);
            accC[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(clo, bh, accC[t], 0, 0, 0);
        }
    }
    // lane (g, il) holds rows i = 4g + r of column il of each 16-column tile
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int t = 0; t < 3; t++) out[w][4 * g + r][16 * t + il] = accC[t][r];
        out[w][4 * g + r][48 + il] = accA[r];
    }
    nr[w][g][il][0] = nacc.x;
    nr[w][g][il][1] = nacc.y;
    nr[w][g][il][2] = nacc.z;
    __syncthreads();
    if (lane >= 16 || li >= n) return;
    float3 f = make_float3(0.f, 0.f, 0.f);
    if (vi) {
        const float *o = out[w][il];
        const float av[3] = {xi.x - cx0, xi.y - cy0, xi.z - cz0}, cp[3] = {Cp.x, Cp.y, Cp.z};
        const float aF = av[0] * o[13] + av[1] * o[14] + av[2] * o[15];
        float fv[3];
#pragma unroll
        for (int al = 0; al < 3; al++) {
            // cii part
            float v = fmaf(cp[al], o[51 + al], o[48 + al]);
            // c d (d . F)
            v += o[al] - (av[0] * o[3 + 3 * al] + av[1] * o[4 + 3 * al] + av[2] * o[5 + 3 * al]) - av[al] * o[12] +
                 av[al] * aF;
            // sum_ga C'_ga c d_al d_ga R_ga
#pragma unroll
            for (int ga = 0; ga < 3; ga++) {
                const int k = 3 * al + ga, kd = 4 * ga;   // r_al R_ga and r_ga R_ga
                const float rR = k < 7 ? o[25 + k] : o[32 + k - 7];
                const float rRd = kd < 7 ? o[25 + kd] : o[32 + kd - 7];
                v = fmaf(cp[ga], o[16 + k] - av[ga] * rR - av[al] * rRd + av[al] * av[ga] * o[34 + ga], v);
            }
            fv[al] = v + nr[w][0][il][al] + nr[w][1][il][al] + nr[w][2][il][al] + nr[w][3][il][al];
        }
        f = make_float3(fv[0], fv[1], fv[2]);
    }
    partial[(size_t)u.w * chunk_stride + bi] = make_float4(f.x, f.y, f.z, 0.f);
}

hipError_t launch_tea_eps(mte_engine *e) {
    if (e->tune.tea_eps == MTE_TEA_EPS_ONEPASS) {
        if (e->periodic)
            hipLaunchKernelGGL(k_tea_eps_sym<true>, dim3(e->tea.eps_nunits), dim3(EPS_GROUP), 0, e->stream, e->P,
                               e->T, e->S[e->cur], e->tea.eps_units, e->tea.eps_ngrp, e->tea.eps_part);
        else
            hipLaunchKernelGGL(k_tea_eps_sym<false>, dim3(e->tea.eps_nunits), dim3(EPS_GROUP), 0, e->stream, e->P,
                               e->T, e->S[e->cur], e->tea.eps_units, e->tea.eps_ngrp, e->tea.eps_part);
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
        hipLaunchKernelGGL(k_tea_pairs_mfma, dim3(e->tea.nunits), dim3(256), 0, e->stream, e->P, e->T, e->S[e->cur],
                           e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->rb_order, e->nb,
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
