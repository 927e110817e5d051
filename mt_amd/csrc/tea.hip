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

// Epsilon partials and ci of bead i over all j of its trajectory
// (integrateTea_epsilon_unlisted, src/bdhitea_kernel.cu:61-101), with the
// block sums in closed form (|e| = 1):
//   sum of the 9 entries of D_hat      = 3 cii + crr (ex + ey + ez)^2
//   row a sum of squares               = cii^2 + crr (2 cii + crr) e_a^2
// Extra beads and the j == i slot enter with weight 0 (their distance kept
// finite), so no branch sits in the pair loop.
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
    const float box = T.tbox[tl.x];
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
            const int self = li - j0;
#pragma unroll 4
            for (int jj = 0; jj < m; jj++) {
                const float4 xj = sx[jj];
                const float3 d = min_image(sub3(f3(xj), ri), box);
                const bool is_self = jj == self;
                const float wj = is_self ? 0.f : xj.w;
                const float r2 = is_self ? 1.f : fmaxf(d.x * d.x + d.y * d.y + d.z * d.z, 1e-12f);
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
    if (active) {
        ci[b] = make_float4(scc + sx2, scc + sy2, scc + sz2, 0.f);
        eps[b] = sall;
    }
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
template <int CHUNK>
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
        const float box = T.tbox[u.x];
        for (int q = 0; q < jn; q++) {
            const float4 xj = sx[q], Fj = sf[q], Rj = sr[q];
            const float3 d = min_image(sub3(f3(xj), ri), box);
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
    const int nc = nchunks[t];
    for (int c = 0; c < nc; c++) {
        const float4 pp = partial[(size_t)c * chunk_stride + b];
        fx += pp.x;
        fy += pp.y;
        fz += pp.z;
    }
    float4 x = Sin.xyz[b], an = Sin.ang[b];
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

hipError_t launch_tea_eps(mte_engine *e) {
    hipLaunchKernelGGL(k_tea_eps, dim3(e->ntiles), dim3(TEA_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                       e->tiles, e->tea.ci, e->tea.eps);
    hipLaunchKernelGGL(k_tea_beta, dim3(e->ntraj), dim3(256), 0, e->stream, e->P, e->T, e->tea.eps, e->tea.beta,
                       e->tea.stats, e->tea.status);
    return hipGetLastError();
}

hipError_t launch_tea_pairs(mte_engine *e) {
    // the kernel's chunk must be the one the work units were built for
    switch (e->tea.unit_chunk) {
    case 64:
        hipLaunchKernelGGL(k_tea_pairs<64>, dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P, e->T,
                           e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb,
                           e->tea.partial);
        break;
    case 256:
        hipLaunchKernelGGL(k_tea_pairs<256>, dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P, e->T,
                           e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb,
                           e->tea.partial);
        break;
    case 128:
        hipLaunchKernelGGL(k_tea_pairs<128>, dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P, e->T,
                           e->S[e->cur], e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb,
                           e->tea.partial);
        break;
    default:
        return hipErrorInvalidValue;
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
