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
// Work layout: one workgroup per (trajectory, 256-bead tile of i), the j range
// of the trajectory streamed through LDS (position + extra flag, force, random
// force) in 256-bead tiles; each (i, j) block costs one rsqrt and ~25 FP32
// operations, so the product is VALU-bound.
#include "engine_internal.h"

namespace mte {

constexpr int TEA_TILE = 256;

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
                const float dx = xj.x - ri.x, dy = xj.y - ri.y, dz = xj.z - ri.z;
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

// Forces for TEA plus the random forces of every bead
// (integrateTea_prepare, src/bdhitea_kernel.cu:16-36: all beads draw).
struct TeaNoiseArgs {
    uint4 *seeds;
    const float *noise;
    float4 *R;
};

__global__ void __launch_bounds__(TEA_TILE)
    k_tea_noise(mte_params P, int nb, TeaNoiseArgs A) {
    const int b = blockIdx.x * TEA_TILE + threadIdx.x;
    if (b >= nb) return;
    const float var = sqrtf(2.0f * KB * P.temp * P.gammaR / P.dt);
    float4 r;
    if (A.noise) {
        const float *nz = A.noise + 6 * (size_t)b;
        r = make_float4(nz[0], nz[1], nz[2], 0.f);
    } else {
        uint4 s = A.seeds[b];
        r = normals4(s);
        A.seeds[b] = s;
    }
    A.R[b] = make_float4(r.x * var, r.y * var, r.z * var, 0.f);
}

// One TEA integration step (integrateTea_kernel_unlisted,
// src/bdhitea_kernel.cu:148-213).
__global__ void __launch_bounds__(TEA_TILE)
    k_tea_step(mte_params P, DevTopo T, DevState Sin, DevState Sout, const float4 *fxyz, const float4 *fang,
               const float4 *R, const float4 *ci, const float *beta, uint4 *seeds, const float *noise,
               const int2 *tiles, LatOffsets L) {
    __shared__ float4 sx[TEA_TILE], sf[TEA_TILE], sr[TEA_TILE];
    const int2 tl = tiles[blockIdx.x];
    const int t = tl.x;
    const int base = T.tbase[t], n = T.tsize[t];
    const int li = tl.y + threadIdx.x;
    const bool active = li < n;
    const int b = base + li;
    const float bt = beta[t];
    float3 ri = make_float3(0, 0, 0), f = ri, c = ri;
    bool skip = true;
    if (active) {
        ri = f3(Sin.xyz[b]);
        const float4 cc = ci[b];
        const float4 Fi = fxyz[b], Ri = R[b];
        const float w = bt * bt;
        const float cx = 1.f / sqrtf(1.f + w * cc.x), cy = 1.f / sqrtf(1.f + w * cc.y),
                    cz = 1.f / sqrtf(1.f + w * cc.z);
        f = make_float3(Fi.x + Ri.x * cx, Fi.y + Ri.y * cy, Fi.z + Ri.z * cz);
        c = make_float3(cx * bt, cy * bt, cz * bt);
        skip = T.extra[b] != 0;
    }
    const float a = P.tea_a;
    for (int j0 = 0; j0 < n; j0 += TEA_TILE) {
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            float4 v = Sin.xyz[base + jl];
            v.w = T.extra[base + jl] ? 1.f : 0.f;
            sx[threadIdx.x] = v;
            sf[threadIdx.x] = fxyz[base + jl];
            sr[threadIdx.x] = R[base + jl];
        }
        __syncthreads();
        const int m = min(TEA_TILE, n - j0);
        if (active && !skip) {
            for (int jj = 0; jj < m; jj++) {
                const float4 xj = sx[jj];
                if (j0 + jj == li || xj.w != 0.f) continue;
                const float4 Fj = sf[jj], Rj = sr[jj];
                const float dx = xj.x - ri.x, dy = xj.y - ri.y, dz = xj.z - ri.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                const float rinv = rsqrtf(r2);
                const float r = r2 * rinv;
                const RpyCoef k = rpy_coef(r, rinv, a);
                const float vx = Fj.x + c.x * Rj.x, vy = Fj.y + c.y * Rj.y, vz = Fj.z + c.z * Rj.z;
                const float s = k.crr * (rinv * rinv) * (dx * vx + dy * vy + dz * vz);
                f.x += k.cii * vx + s * dx;
                f.y += k.cii * vy + s * dy;
                f.z += k.cii * vz + s * dz;
            }
        }
        __syncthreads();
    }
    if (!active) return;
    float4 x = Sin.xyz[b], an = Sin.ang[b];
    const float mult = P.dt / P.gammaR;
    const float nx = ri.x + mult * f.x, ny = ri.y + mult * f.y, nz = ri.z + mult * f.z;
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
        x.x = nx;
        x.y = ny;
        x.z = nz;
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

hipError_t launch_tea_step(mte_engine *e, const float *noise_step) {
    TeaNoiseArgs A{e->seeds, noise_step, e->tea.R};
    hipLaunchKernelGGL(k_tea_noise, dim3((e->nb + TEA_TILE - 1) / TEA_TILE), dim3(TEA_TILE), 0, e->stream, e->P,
                       e->nb, A);
    hipLaunchKernelGGL(k_tea_step, dim3(e->ntiles), dim3(TEA_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                       e->S[1 - e->cur], e->fxyz, e->fang, e->tea.R, e->tea.ci, e->tea.beta, e->seeds, noise_step,
                       e->tiles, e->L);
    e->cur = 1 - e->cur;
    return hipGetLastError();
}

}  // namespace mte
