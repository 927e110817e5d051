// device_common.h -- device-side building blocks shared by the gfx950 kernels.
//
// Bonded geometry is written in attachment-point form: each bonded gap is
// g = (r_i + a_i) - (r_j + b_j) with body-fixed offsets rotated by
// R = Rz(psi) Ry(theta) Rx(phi) (the rotation whose entries the reference
// expands at src/compute_cuda.cu:89-96 and :337-352).  The gradient of
// |g|^2/2 with respect to bead i is g for x, y, z and u_q . (a_i x g) for the
// angles, u_phi = Rz Ry e_x, u_theta = Rz e_y, u_psi = e_z.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mt_engine.h"

namespace mte {

constexpr float RMON = 2.0f;          // src/mt.h:24
constexpr float LJ_CUT = 6.0f;        // src/mt.h:25
constexpr float ANGLE_CUT = 1.0f;     // src/mt.h:42
constexpr float PAIR_CUT = 2.5f;      // src/mt.h:61
constexpr float KB = 0.0019872041f;   // src/mt.h:40
constexpr int MAXLON = 8;             // src/preparator.cpp:722
constexpr int MAXLAT = 16;            // src/preparator.cpp:721
// Centre distance beyond which no bonded candidate can pass PAIR_CUT:
// |g| >= d - 2*r_mon (axial) and |g| >= d - |p1| - |p2| ~ d - 3.995 (lateral).
constexpr float BOND_PREFILTER = 6.75f;

// Lateral attachment offsets p1, p2 of src/mt.h:44-50, evaluated in double on
// the host and rounded once.
struct LatOffsets { float p1x, p1y, p1z, p2x, p2y, p2z; };

// Packed list entries: (global j << 1) | flag, where flag is "R_MON > 0" for
// harmonic/longitudinal entries and "p1/p2 swapped" for lateral entries.
__device__ __forceinline__ int ent_j(int e) { return e >> 1; }
__device__ __forceinline__ int ent_f(int e) { return e & 1; }

// Topology on the device (global bead indices, k-major list slots).
struct DevTopo {
    int nb, ntraj, max_lj;
    const int *tid;       // bead -> trajectory
    const int *tbase;     // trajectory -> first bead
    const int *tsize;     // trajectory -> bead count
    const int *harm;      // packed partner or -1
    const int *harm_raw;  // reference-encoded partner (local), drives pairs_kernel semantics
    int *lon_n, *lon;     // lon[k * nb + b]
    int *lat_n, *lat;
    int *lj_n, *lj;
    const uint8_t *fixed;
    uint8_t *extra;
    const int *mon_type;
    int *gtp, *on_tub;
    const float *zs;
    const float *tbox;    // trajectory -> periodic box side (0: open boundaries)
    int *overflow;        // set to 1 by a rebuild that ran out of list capacity
    int *frozen;          // steps the resident kernel left undone because `overflow` was set
};

// One buffer of bead state: positions, angles and the rotated offsets.
struct DevState {
    float4 *xyz;   // x, y, z, -
    float4 *ang;   // fi, theta, psi, -   (Coord order)
    float4 *fn;    // R e_z
    float4 *fp1;   // R p1
    float4 *fp2;   // R p2
};

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }

// Minimum-image separation in a cubic periodic box of side L (L == 0: open).
// The reference has no periodic boundaries; the uneven free-dimer batches
// run in one.
__device__ __forceinline__ float3 min_image(float3 d, float L) {
    if (L > 0.f) {
        const float iL = 1.f / L;
        d.x -= L * rintf(d.x * iL);
        d.y -= L * rintf(d.y * iL);
        d.z -= L * rintf(d.z * iL);
    }
    return d;
}
// min_image with 1/L taken outside a loop (iL = 0 on open boundaries, where
// d comes back unchanged): no branch in the loop body
__device__ __forceinline__ float3 min_image_inv(float3 d, float L, float iL) {
    d.x -= L * rintf(d.x * iL);
    d.y -= L * rintf(d.y * iL);
    d.z -= L * rintf(d.z * iL);
    return d;
}
__device__ __forceinline__ float inv_side(float L) { return L > 0.f ? 1.f / L : 0.f; }
// min_image with iL = inv_side(L) taken by the caller (outside its loops):
// the same values, without a division per call
__device__ __forceinline__ float3 min_image_pre(float3 d, float L, float iL) {
    if (L > 0.f) d = min_image_inv(d, L, iL);
    return d;
}
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }

// Periodic cell index: i in [-dims, 2 dims) wrapped into [0, dims) (a
// neighbour stencil never reaches further), without an integer division.
__device__ __forceinline__ int wrap_cell(int i, int dims) { return i < 0 ? i + dims : (i >= dims ? i - dims : i); }

// sin and cos by Cody-Waite reduction modulo pi/2 (two-part pi/2, exact
// with FMA for |x| < 8192) and the single-precision minimax polynomials of
// the Cephes library (S. L. Moshier, sinf.c / cosf.c) on [-pi/4, pi/4];
// about 1 ulp.  |x| >= 8192 is first reduced modulo 2*pi in FP64.
__device__ __forceinline__ void sincos_fast(float x, float *s, float *c) {
    if (!(fabsf(x) < 8192.f)) {
        const double d = (double)x;
        x = (float)(d - 6.283185307179586477 * rint(d * 0.15915494309189533577));
    }
    const float j = rintf(x * 0.63661977236758134f);
    const int q = (int)j;
    float r = fmaf(j, -1.57079637050628662109375f, x);
    r = fmaf(j, 4.37113900018624283e-8f, r);
    const float z = r * r;
    const float sp = fmaf(z * r, fmaf(z, fmaf(z, -1.9515295891e-4f, 8.3321608736e-3f), -1.6666654611e-1f), r);
    const float cp = fmaf(z * z, fmaf(z, fmaf(z, 2.443315711809948e-5f, -1.388731625493765e-3f), 4.166664568298827e-2f),
                          fmaf(-0.5f, z, 1.0f));
    float ss = (q & 1) ? cp : sp;
    float cc = (q & 1) ? sp : cp;
    if (q & 2) ss = -ss;
    if ((q + 1) & 2) cc = -cc;
    *s = ss;
    *c = cc;
}
__device__ __forceinline__ float sin_fast(float x) {
    float s, c;
    sincos_fast(x, &s, &c);
    return s;
}
// sin_fast for arguments that usually need no reduction (bonded angle
// differences): when x rounds to quadrant 0 the reduction is the identity
// and only the sine polynomial is evaluated -- the same value sin_fast
// returns -- and a wave takes the full path only if one of its lanes needs it.
__device__ __forceinline__ float sin_fast_near0(float x) {
    if (rintf(x * 0.63661977236758134f) == 0.f) {
        const float z = x * x;
        return fmaf(z * x, fmaf(z, fmaf(z, -1.9515295891e-4f, 8.3321608736e-3f), -1.6666654611e-1f), x);
    }
    return sin_fast(x);
}
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_fast(float x) { return __builtin_amdgcn_rsqf(x); }

struct Rot {
    float cf, sf, ct, st, cp, sp;
    __device__ __forceinline__ float3 apply(float vx, float vy, float vz) const {
        return make_float3(cp * ct * vx + (cp * st * sf - sp * cf) * vy + (cp * st * cf + sp * sf) * vz,
                           sp * ct * vx + (sp * st * sf + cp * cf) * vy + (sp * st * cf - cp * sf) * vz,
                           -st * vx + ct * sf * vy + ct * cf * vz);
    }
};

__device__ __forceinline__ Rot make_rot(float fi, float theta, float psi) {
    Rot r;
    sincos_fast(fi, &r.sf, &r.cf);
    sincos_fast(theta, &r.st, &r.ct);
    sincos_fast(psi, &r.sp, &r.cp);
    return r;
}

// Rotated offsets of one bead.
__device__ __forceinline__ void frame_vectors(float fi, float theta, float psi, const LatOffsets &L, float4 &n,
                                              float4 &a1, float4 &a2) {
    Rot R = make_rot(fi, theta, psi);
    float3 v = R.apply(0.f, 0.f, 1.f);
    n = make_float4(v.x, v.y, v.z, 0.f);
    v = R.apply(L.p1x, L.p1y, L.p1z);
    a1 = make_float4(v.x, v.y, v.z, 0.f);
    v = R.apply(L.p2x, L.p2y, L.p2z);
    a2 = make_float4(v.x, v.y, v.z, 0.f);
}

// Gap g = dij + (a - b), dij = r_i - r_j (minimum image), and the gradient
// of |g|^2/2 with respect to bead i.  grad order: x, y, z, fi, theta, psi.
// uphi = (cp*ct, sp*ct, -st), uth = (-sp, cp, 0).
struct Grad { float g[6]; float dr, inv; };   // inv = 1/dr (0 when dr == 0)

__device__ __forceinline__ Grad gap_grad(float3 dij, float3 a, float3 b, float3 uphi, float2 uth) {
    Grad o;
    const float gx = dij.x + (a.x - b.x);
    const float gy = dij.y + (a.y - b.y);
    const float gz = dij.z + (a.z - b.z);
    const float t0 = a.y * gz - a.z * gy;
    const float t1 = a.z * gx - a.x * gz;
    const float t2 = a.x * gy - a.y * gx;
    o.g[0] = gx; o.g[1] = gy; o.g[2] = gz;
    o.g[3] = uphi.x * t0 + uphi.y * t1 + uphi.z * t2;
    o.g[4] = uth.x * t0 + uth.y * t1;
    o.g[5] = t2;
    const float d2 = gx * gx + gy * gy + gz * gz;
    o.inv = d2 > 0.f ? rsq_fast(d2) : 0.f;
    o.dr = d2 * o.inv;
    return o;
}

// src/compute_cuda.cu:16-30
__device__ __forceinline__ float dmorse(float D, float a, float x) {
    const float e = __expf(-a * x);
    return 2.f * a * D * (1.f - e) * e;
}
__device__ __forceinline__ float morse_en(float D, float a, float x) {
    const float e = 1.f - __expf(-a * x);
    return D * e * e - D;
}
__device__ __forceinline__ float dbarr(float a, float r, float w, float x) {
    const float iw2 = rcp_fast(w * w);
    return -a * __expf(-0.5f * (x - r) * (x - r) * iw2) * (x - r) * iw2;
}
__device__ __forceinline__ float barr(float a, float r, float w, float x) {
    return a * __expf(-0.5f * (x - r) * (x - r) * rcp_fast(w * w));
}

// ---------------------------------------------------------------- HybridTaus
// GPU Gems 3, ch. 37 combined Tausworthe + LCG generator with the constants of
// src/HybridTaus.cu:72-82 and src/ht.cuh:8-10.
__device__ __forceinline__ uint32_t taus_step(uint32_t &z, int s1, int s2, int s3, uint32_t m) {
    const uint32_t b = (((z << s1) ^ z) >> s2);
    return z = (((z & m) << s3) ^ b);
}
__device__ __forceinline__ uint32_t hybrid_taus(uint4 &s) {
    return taus_step(s.x, 13, 19, 12, 4294967294u) ^ taus_step(s.y, 2, 25, 4, 4294967288u) ^
           taus_step(s.z, 3, 11, 17, 4294967280u) ^ (s.w = 1664525u * s.w + 1013904223u);
}
__device__ __forceinline__ float u01(uint32_t u) {
    float f = __uint_as_float(0x3f800000u | (0x007fffffu & u)) - 1.0f;
    return f == 0.f ? 1.0e-8f : f;
}
// Box-Muller radius sqrt(-2 ln u) = sqrt(-2 ln2 log2 u) from the hardware
// log2 and square root (1 ulp each).  u lies in [2^-23, 1) or is 1e-8, so
// both arguments are normal floats and need none of the denormal scaling and
// rounding fix-ups that logf and sqrtf expand to.
__device__ __forceinline__ float bm_radius(float u) {
    return __builtin_amdgcn_sqrtf(-1.38629436111989061883f * __builtin_amdgcn_logf(u));
}
// One draw of four normals (two Box-Muller pairs), advancing the state four
// times as src/HybridTaus.cu:85-98 does.
__device__ __forceinline__ float4 normals4(uint4 &s) {
    float4 o;
    float r = bm_radius(u01(hybrid_taus(s)));
    float th = 6.28318530717958647692f * u01(hybrid_taus(s));
    o.x = r * __sinf(th);
    o.y = r * __cosf(th);
    r = bm_radius(u01(hybrid_taus(s)));
    th = 6.28318530717958647692f * u01(hybrid_taus(s));
    o.z = r * __sinf(th);
    o.w = r * __cosf(th);
    return o;
}

}  // namespace mte
