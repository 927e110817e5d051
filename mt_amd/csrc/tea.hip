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
// group, 128-bead j chunk), the chunk's positions, forces and random forces in
// LDS, FP32 partial sums per chunk reduced in fixed order by the finish kernel,
// which adds the self term and integrates.  Each (i, j) block costs one rsqrt
// and ~35 FP32 operations: the product is VALU-bound.
#include "engine_internal.h"

namespace mte {

constexpr int TEA_TILE = 256;
constexpr int TEA_CHUNK = TEA_CHUNK_BEADS;

struct RpyCoef { float crr, cii; };

// D_hat coefficients for distance r (src/bdhitea_kernel.cu:38-58):
// D_hat = cii I + crr e e^T.
__device__ __forceinline__ RpyCoef rpy_coef(float r, float rinv, float a) {
    RpyCoef c;
    const float ra = r / a;
    if (ra > 2.f) {
        const float inv = 1.f / ra;
        const float inv2 = inv * inv;
        c.crr = 0.75f * inv * (1.f - 2.f * inv2);
        c.cii = 0.75f * inv * (1.f + (2.f / 3.f) * inv2);
    } else {
        c.crr = 3.f * ra / 32.f;
        c.cii = 1.f - 9.f * ra / 32.f;
    }
    return c;
}

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
    float sxx = 0.f, syy = 0.f, szz = 0.f, sall = 0.f;
    for (int j0 = 0; j0 < n; j0 += TEA_TILE) {
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            float4 v = S.xyz[base + jl];
            v.w = T.extra[base + jl] ? 1.f : 0.f;
            sx[threadIdx.x] = v;
        }
        __syncthreads();
        const int m = min(TEA_TILE, n - j0);
        if (active && !skip) {
            for (int jj = 0; jj < m; jj++) {
                const float4 xj = sx[jj];
                if (j0 + jj == li || xj.w != 0.f) continue;
                const float3 d = min_image(sub3(f3(xj), ri), box);
                const float dx = d.x, dy = d.y, dz = d.z;
                const float r = sqrtf(dx * dx + dy * dy + dz * dz);
                const float ex = dx / r, ey = dy / r, ez = dz / r;
                const RpyCoef c = rpy_coef(r, 1.f / r, P.tea_a);
                const float dxx = ex * ex * c.crr + c.cii, dxy = ex * ey * c.crr, dxz = ex * ez * c.crr;
                const float dyy = ey * ey * c.crr + c.cii, dyz = ey * ez * c.crr, dzz = ez * ez * c.crr + c.cii;
                sall += dxx + 2.f * dxy + 2.f * dxz + dyy + 2.f * dyz + dzz;
                sxx += dxx * dxx + dxy * dxy + dxz * dxz;
                syy += dxy * dxy + dyy * dyy + dyz * dyz;
                szz += dxz * dxz + dyz * dyz + dzz * dzz;
            }
        }
        __syncthreads();
    }
    if (active) {
        ci[b] = make_float4(sxx, syy, szz, 0.f);
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
// noise at a far position, so they contribute nothing and need no branch.
// Splitting j into chunks gives ~10 work units per SIMD at 64 x 1040 beads,
// against ~1 per SIMD for one thread per bead over the whole trajectory.
__global__ void __launch_bounds__(TEA_TILE)
    k_tea_pairs(mte_params P, DevTopo T, DevState S, const float4 *fxyz, const float4 *R, const float4 *ci,
                const float *beta, const int4 *units, int chunk_stride, float4 *partial) {
    __shared__ float4 sx[TEA_CHUNK], sf[TEA_CHUNK], sr[TEA_CHUNK];
    const int4 u = units[blockIdx.x];   // traj, i0, j0, chunk index
    const int base = T.tbase[u.x], n = T.tsize[u.x];
    const int jn = min(TEA_CHUNK, n - u.z);
    for (int q = threadIdx.x; q < jn; q += TEA_TILE) {
        const int jb = base + u.z + q;
        float4 x = S.xyz[jb];
        float4 f = fxyz[jb], r = R[jb];
        if (T.extra[jb]) {
            x = make_float4(3.0e18f, 3.0e18f, 3.0e18f, 0.f);
            f = r = make_float4(0.f, 0.f, 0.f, 0.f);
        }
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
        const int self = li - u.z;   // position of i in this chunk, if any
        for (int q = 0; q < jn; q++) {
            const float4 xj = sx[q], Fj = sf[q], Rj = sr[q];
            const float3 d = min_image(sub3(f3(xj), ri), box);
            const float dx = d.x, dy = d.y, dz = d.z;
            const bool is_self = q == self;
            const float r2 = is_self ? 1.f : dx * dx + dy * dy + dz * dz;   // keeps the j == i slot finite
            const float rinv = rsq_fast(r2);
            const float r = r2 * rinv;
            // D_hat = cii I + crr e e^T (src/bdhitea_kernel.cu:38-58); both branches, then select
            const float u1 = a * rinv, u2 = u1 * u1;
            const float far_ii = 0.75f * u1 * fmaf(0.66666666666666667f, u2, 1.f);
            const float far_rr = 0.75f * u1 * fmaf(-2.f, u2, 1.f);
            const float ra = r * ia;
            const bool near = !(ra > 2.f);
            float cii = near ? fmaf(-0.28125f, ra, 1.f) : far_ii;
            float crr = near ? 0.09375f * ra : far_rr;
            if (is_self) cii = crr = 0.f;
            const float vx = fmaf(cx, Rj.x, Fj.x), vy = fmaf(cy, Rj.y, Fj.y), vz = fmaf(cz, Rj.z, Fj.z);
            const float s = (crr * rinv * rinv) * (dx * vx + dy * vy + dz * vz);
            acc.x += fmaf(cii, vx, s * dx);
            acc.y += fmaf(cii, vy, s * dy);
            acc.z += fmaf(cii, vz, s * dz);
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
    hipLaunchKernelGGL(k_tea_pairs, dim3(e->tea.nunits), dim3(TEA_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                       e->fxyz, e->tea.R, e->tea.ci, e->tea.beta, e->tea.units, e->nb, e->tea.partial);
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
