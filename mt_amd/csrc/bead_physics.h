// bead_physics.h -- per-bead force and energy evaluation, shared by the
// streamed kernels (state in HBM) and the resident kernel (state in LDS).
// An accessor type supplies neighbour data by global bead index.
//
// Restates compute_kernel (src/compute_cuda.cu:32-525) and energy_kernel
// (:676-911) in attachment-point form; see device_common.h.
#pragma once
#include "device_common.h"

namespace mte {

__device__ __forceinline__ float3 scale3(float s, float3 v) { return make_float3(s * v.x, s * v.y, s * v.z); }

// Own-bead quantities needed for the angular gradients.
struct Own {
    float3 r, n, a1, a2, uphi;
    float2 uth;
    float fi, theta, psi;
    float box;   // periodic side of the bead's trajectory (0: open)
    float ibox;  // inv_side(box)
};
__device__ __forceinline__ void set_box(Own &o, float L) {
    o.box = L;
    o.ibox = inv_side(L);
}

// trig = (sin theta, cos theta, sin psi, cos psi) of the bead.
__device__ __forceinline__ Own make_own_trig(float4 xyz, float4 ang, float4 fn, float4 fp1, float4 fp2, float4 trig) {
    Own o;
    o.r = f3(xyz);
    o.n = f3(fn);
    o.a1 = f3(fp1);
    o.a2 = f3(fp2);
    o.fi = ang.x;
    o.theta = ang.y;
    o.psi = ang.z;
    o.uphi = make_float3(trig.w * trig.y, trig.z * trig.y, -trig.x);
    o.uth = make_float2(-trig.z, trig.w);
    o.box = o.ibox = 0.f;
    return o;
}

__device__ __forceinline__ Own make_own(float4 xyz, float4 ang, float4 fn, float4 fp1, float4 fp2) {
    Own o;
    o.r = f3(xyz);
    o.n = f3(fn);
    o.a1 = f3(fp1);
    o.a2 = f3(fp2);
    o.fi = ang.x;
    o.theta = ang.y;
    o.psi = ang.z;
    float st, ct, sp, cp;
    sincos_fast(ang.y, &st, &ct);
    sincos_fast(ang.z, &sp, &cp);
    o.uphi = make_float3(cp * ct, sp * ct, -st);
    o.uth = make_float2(-sp, cp);
    o.box = o.ibox = 0.f;
    return o;
}

// Bending torques of a harmonic or longitudinal pair (:156-177, :271-295).
__device__ __forceinline__ void bending(const mte_params &P, float rm, float th0, const Own &I, float4 aj,
                                        float F[6]) {
    const float psiji = aj.z - I.psi, thetaji = aj.y - I.theta, fiji = aj.x - I.fi;
    // rm > 0: F += B sin(d - d0); rm < 0: F -= B sin(-d - d0).  One branch-free
    // form with s = +-1 (exact sign flips: the same values), since the two
    // monomers of a dimer sit in neighbouring lanes and take opposite signs.
    const float sg = rm > 0.f ? 1.f : -1.f;
    F[5] += sg * P.B_psi * sin_fast_near0(sg * psiji - P.psi_0);
    F[3] += sg * P.B_fi * sin_fast_near0(sg * fiji - P.fi_0);
    F[4] += sg * P.B_theta * sin_fast_near0(sg * thetaji - th0);
}

// Per-bead flags packed into xyz.w when a trajectory's state is staged in LDS.
constexpr int FL_MON = 1, FL_GTP1 = 2, FL_OFFTUB = 4;

__device__ __forceinline__ int pack_flags(const DevTopo &T, int b) {
    return (T.mon_type[b] ? FL_MON : 0) | (T.gtp[b] == 1 ? FL_GTP1 : 0) | (T.on_tub[b] == 0 ? FL_OFFTUB : 0);
}

// A trajectory's bead data staged in LDS, by local index: position with the
// packed flags in w, and the rotated offsets with the Euler angles in w
// (n|fi, p1|theta, p2|psi).
struct LdsAcc {
    const float4 *x, *n, *p1, *p2;
    __device__ __forceinline__ float4 xyz(int j) const { return x[j]; }
    __device__ __forceinline__ float4 ang(int j) const { return make_float4(n[j].w, p1[j].w, p2[j].w, 0.f); }
    __device__ __forceinline__ float4 fn(int j) const { return n[j]; }
    __device__ __forceinline__ float4 fp1(int j) const { return p1[j]; }
    __device__ __forceinline__ float4 fp2(int j) const { return p2[j]; }
    __device__ __forceinline__ int flags(int j) const { return __float_as_int(x[j].w); }
    __device__ __forceinline__ int mon(int j) const { return flags(j) & FL_MON; }
    __device__ __forceinline__ bool gtp1(int j) const { return (flags(j) & FL_GTP1) != 0; }
    __device__ __forceinline__ bool offtub(int j) const { return (flags(j) & FL_OFFTUB) != 0; }
};

// Stages beads [0, n) of the trajectory starting at global bead `base` into
// the four LDS arrays (all threads of the workgroup; no barrier).
__device__ __forceinline__ void stage_state(const DevTopo &T, const DevState &S, int base, int n, float4 *lx,
                                            float4 *ln, float4 *l1, float4 *l2) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int b = base + i;
        float4 x = S.xyz[b];
        x.w = __int_as_float(pack_flags(T, b));
        const float4 a = S.ang[b];
        float4 fn = S.fn[b], f1 = S.fp1[b], f2 = S.fp2[b];
        fn.w = a.x;
        f1.w = a.y;
        f2.w = a.z;
        lx[i] = x;
        ln[i] = fn;
        l1[i] = f1;
        l2[i] = f2;
    }
}

// Neighbour lists held in the k-major global arrays of bead b, with entry
// indices taken relative to `base` (0: global indices, the streamed kernels;
// the trajectory's first bead: local indices, the resident kernel).
struct GlobalLists {
    const DevTopo &T;
    int b, base;
    __device__ __forceinline__ int harm() const {
        const int h = T.harm[b];
        return h >= 0 ? h - 2 * base : -1;
    }
    __device__ __forceinline__ int nlon() const { return T.lon_n[b]; }
    __device__ __forceinline__ int lon(int k) const { return T.lon[k * T.nb + b] - 2 * base; }
    __device__ __forceinline__ int nlat() const { return T.lat_n[b]; }
    __device__ __forceinline__ int lat(int k) const { return T.lat[k * T.nb + b] - 2 * base; }
    __device__ __forceinline__ int nlj() const { return T.lj_n[b]; }
    __device__ __forceinline__ int lj(int k) const { return T.lj[k * T.nb + b] - base; }
};

// Position of a bead's k-th LJ neighbour; a list format that stores
// something other than the neighbour's index provides its own overload.
template <class Acc, class Lists>
__device__ __forceinline__ float4 lj_xyz(const Acc &A, const Lists &Ls, int k) {
    return A.xyz(Ls.lj(k));
}

// LJ repulsion over the list; the cutoff test selects the force (no branch
// around the accumulation), and periodic and open boxes get separate loops,
// as do runs with and without the regularization cap (a uniform choice: the
// cap's reciprocal square root is not paid per pair when it is off).
template <bool PERIODIC, bool REG, class Acc, class Lists>
__device__ __forceinline__ void lj_sum(const mte_params &P, const Acc &A, const Lists &Ls, const Own &I, float F[6]) {
    const int nlj = Ls.nlj();
    const float s6 = 6.f * (P.ljscale * P.ljsigma6);
    // REGULARIZATION (src/compute_cuda.cu:482-486): the pair force s r is
    // capped at gammaR r_mon / dt, i.e. s at that over r (r_mon = 2 nm)
    const float cap = P.gammaR * 2.f / P.dt;
    float fx = F[0], fy = F[1], fz = F[2];   // same accumulation order, scalar registers
    auto pair = [&](float4 xj) {
        float3 d = sub3(I.r, f3(xj));
        if (PERIODIC) d = min_image_inv(d, I.box, I.ibox);   // called with I.box > 0
        const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
        const float inv = rcp_fast(r2);
        const float inv2 = inv * inv;
        float s = r2 < LJ_CUT * LJ_CUT ? s6 * (inv2 * inv2) : 0.f;
        if (REG) s = fminf(s, cap * rsq_fast(r2));
        fx += s * d.x;
        fy += s * d.y;
        fz += s * d.z;
    };
    // four entries and their positions requested before the first is used
    // (one wait for the four gathers rather than one each), then summed in
    // list order
    int k = 0;
    for (; k + 4 <= nlj; k += 4) {
        const float4 x0 = lj_xyz(A, Ls, k), x1 = lj_xyz(A, Ls, k + 1), x2 = lj_xyz(A, Ls, k + 2),
                     x3 = lj_xyz(A, Ls, k + 3);
        pair(x0);
        pair(x1);
        pair(x2);
        pair(x3);
    }
    for (; k < nlj; k++) pair(lj_xyz(A, Ls, k));
    F[0] = fx;
    F[1] = fy;
    F[2] = fz;
}
template <bool PERIODIC, class Acc, class Lists>
__device__ __forceinline__ void lj_forces(const mte_params &P, const Acc &A, const Lists &Ls, const Own &I,
                                          float F[6]) {
    if (P.lj_regularize) lj_sum<PERIODIC, true>(P, A, Ls, I, F);
    else lj_sum<PERIODIC, false>(P, A, Ls, I, F);
}

// Force and torque on bead b (an accessor index, as are the list entries);
// wall_z is Parameters.zs of the bead's trajectory.
template <class Acc, class Lists>
__device__ __forceinline__ void bead_force(const mte_params &P, const Acc &A, const Lists &Ls, int b, float wall_z,
                                           const Own &I, float F[6]) {
#pragma unroll
    for (int q = 0; q < 6; q++) F[q] = 0.f;
    // harmonic (intra-dimer spring + bending), :70-182
    const int h = Ls.harm();
    if (h >= 0) {
        const int j = ent_j(h);
        const float rm = ent_f(h) ? RMON : -RMON;
        const float3 nj = f3(A.fn(j));
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(A.xyz(j))), I.box, I.ibox), scale3(rm, I.n), scale3(-rm, nj), I.uphi, I.uth);
#pragma unroll
        for (int q = 0; q < 6; q++) F[q] -= P.C * G.g[q];
        if (G.dr < ANGLE_CUT) {
            const float th0 = A.gtp1(b) ? P.theta0_gtp : P.theta0_gdp;
            bending(P, rm, th0, I, A.ang(j), F);
        }
    }
    // longitudinal Morse + barrier + bending, :189-299
    const int nlon = Ls.nlon();
    for (int k = 0; k < nlon; k++) {
        const int e = Ls.lon(k);
        const int j = ent_j(e);
        const float rm = ent_f(e) ? RMON : -RMON;
        const float4 xj = A.xyz(j);
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(xj)), I.box, I.ibox), scale3(rm, I.n), scale3(-rm, f3(A.fn(j))), I.uphi, I.uth);
        float dUdr = (G.dr == 0.f) ? 0.f : dmorse(P.D_long, P.A_long, G.dr) * G.inv;
        if (P.barrier && A.offtub(b) && A.offtub(j) && G.dr != 0.f)
            dUdr += dbarr(P.a_barr_long, P.r_barr_long, P.w_barr_long, G.dr) * G.inv;
#pragma unroll
        for (int q = 0; q < 6; q++) F[q] -= dUdr * G.g[q];
        if (G.dr < ANGLE_CUT) {
            const int last = (I.r.z > xj.z) ? b : j;
            const float th0 = A.gtp1(last) ? P.theta0_gtp : P.theta0_gdp;
            bending(P, rm, th0, I, A.ang(j), F);
        }
    }
    // lateral Morse + barrier with seam weakening, :304-466
    const int nlat = Ls.nlat();
    for (int k = 0; k < nlat; k++) {
        const int e = Ls.lat(k);
        const int j = ent_j(e);
        const int sw = ent_f(e);
        const float3 a = sw ? I.a1 : I.a2;
        const float3 bj = f3(sw ? A.fp2(j) : A.fp1(j));
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(A.xyz(j))), I.box, I.ibox), a, bj, I.uphi, I.uth);
        // seam bonds (monomer types differ) are weakened by seam_coeff (:447-452)
        const float Dl = A.mon(b) != A.mon(j) ? P.D_lat / P.seam_coeff : P.D_lat;
        float dUdr = (G.dr == 0.f) ? 0.f : dmorse(Dl, P.A_lat, G.dr) * G.inv;
        if (P.barrier && A.offtub(b) && A.offtub(j) && G.dr != 0.f)
            dUdr += dbarr(P.a_barr_lat, P.r_barr_lat, P.w_barr_lat, G.dr) * G.inv;
#pragma unroll
        for (int q = 0; q < 6; q++) F[q] -= dUdr * G.g[q];
    }
    // LJ repulsion r^-6 below the hard 6 nm cutoff, :470-495
    if (P.lj_on) {
        if (I.box > 0.f) lj_forces<true>(P, A, Ls, I, F);
        else lj_forces<false>(P, A, Ls, I, F);
    }
    // cylindrical and end walls, :497-517
    if (P.is_wall) {
        const float z = I.r.z, lb = P.rep_leftborder, zt = wall_z + lb;
        if (z < lb) F[2] += P.rep_eps * fabsf(z - lb);
        else if (z > zt) F[2] += -P.rep_eps * fabsf(z - zt);
        const float rad2 = I.r.x * I.r.x + I.r.y * I.r.y;
        if (rad2 > P.rep_r * P.rep_r) {
            const float rad = sqrtf(rad2);
            const float coeff = -P.rep_eps * (rad - P.rep_r);
            F[0] += I.r.x / rad * coeff;
            F[1] += I.r.y / rad * coeff;
        }
    }
}

// Energies in the energy_kernel convention (:676-911): U_harm, U_long,
// U_lat, U_psi, U_fi, U_teta, U_lj, each halved.
template <class Acc>
__device__ __forceinline__ void bead_energy(const mte_params &P, const DevTopo &T, const Acc &A, int b, const Own &I,
                                            float U[7]) {
#pragma unroll
    for (int q = 0; q < 7; q++) U[q] = 0.f;
    const int nb = T.nb;
    const int h = T.harm[b];
    if (h >= 0) {
        const int j = ent_j(h);
        const float rm = ent_f(h) ? RMON : -RMON;
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(A.xyz(j))), I.box, I.ibox), scale3(rm, I.n), scale3(-rm, f3(A.fn(j))), I.uphi, I.uth);
        U[0] += (P.C / 2.f) * G.dr * G.dr;
        if (G.dr < ANGLE_CUT) {
            const float4 aj = A.ang(j);
            const float psiij = I.psi - aj.z, fiij = I.fi - aj.x, thetaji = aj.y - I.theta;
            U[3] += P.B_psi * (1.f - cosf(psiij - P.psi_0));
            U[4] += P.B_fi * (1.f - cosf(fiij - P.fi_0));
            U[5] += P.B_theta * (1.f - cosf(thetaji - (A.gtp1(b) ? P.theta0_gtp : P.theta0_gdp)));
        }
    }
    const int nlon = T.lon_n[b];
    for (int k = 0; k < nlon; k++) {
        const int e = T.lon[k * nb + b];
        const int j = ent_j(e);
        const float rm = ent_f(e) ? RMON : -RMON;
        const float4 xj = A.xyz(j);
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(xj)), I.box, I.ibox), scale3(rm, I.n), scale3(-rm, f3(A.fn(j))), I.uphi, I.uth);
        U[1] += morse_en(P.D_long, P.A_long, G.dr);
        if (P.barrier && A.offtub(b) && A.offtub(j))
            U[1] += barr(P.a_barr_long, P.r_barr_long, P.w_barr_long, G.dr);
        if (G.dr < ANGLE_CUT) {
            const float4 aj = A.ang(j);
            const float psiij = I.psi - aj.z, fiij = I.fi - aj.x, thetaij = I.theta - aj.y;
            const int last = (I.r.z > xj.z) ? b : j;
            const float th0 = A.gtp1(last) ? P.theta0_gtp : P.theta0_gdp;
            U[3] += P.B_psi * (1.f - cosf(psiij - P.psi_0));
            U[4] += P.B_fi * (1.f - cosf(fiij - P.fi_0));
            U[5] += P.B_theta * (1.f - cosf(thetaij - th0));
        }
    }
    const int nlat = T.lat_n[b];
    for (int k = 0; k < nlat; k++) {
        const int e = T.lat[k * nb + b];
        const int j = ent_j(e);
        const int sw = ent_f(e);
        const Grad G = gap_grad(min_image_pre(sub3(I.r, f3(A.xyz(j))), I.box, I.ibox), sw ? I.a1 : I.a2, f3(sw ? A.fp2(j) : A.fp1(j)),
                                I.uphi, I.uth);
        if (A.mon(b) != A.mon(j)) U[2] += morse_en(P.D_lat / P.seam_coeff, P.A_lat, G.dr);
        else U[2] += morse_en(P.D_lat, P.A_lat, G.dr);
        if (P.barrier && A.offtub(b) && A.offtub(j))
            U[2] += barr(P.a_barr_lat, P.r_barr_lat, P.w_barr_lat, G.dr);
    }
    if (P.lj_on) {
        const int nlj = T.lj_n[b];
        for (int k = 0; k < nlj; k++) {
            const int j = T.lj[k * nb + b];
            const float3 d = min_image_pre(sub3(I.r, f3(A.xyz(j))), I.box, I.ibox);
            const float dx = d.x, dy = d.y, dz = d.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (sqrtf(r2) < LJ_CUT) {
                const float inv = 1.f / r2;
                U[6] += P.ljscale * P.ljsigma6 * (inv * inv * inv);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 7; q++) U[q] *= 0.5f;
}

// Euler-Maruyama update of one mobile bead, src/compute_cuda.cu:955-961.
__device__ __forceinline__ void em_update(const mte_params &P, const float F[6], float4 rx, float4 ra, float4 &x,
                                          float4 &a) {
    const float dtg = P.dt / P.gammaR;
    const float dta = P.dt / (P.gammaTheta * P.alpha);
    const float dtt = P.dt / P.gammaTheta;
    const float vta = P.varTheta * sqrtf(P.freeze_temp / P.alpha);
    x.x += dtg * F[0] + P.varR * rx.x;
    x.y += dtg * F[1] + P.varR * rx.y;
    x.z += dtg * F[2] + P.varR * rx.z;
    a.x += dta * F[3] + vta * ra.x;          // fi
    a.z += dta * F[5] + vta * ra.y;          // psi
    a.y += dtt * F[4] + P.varTheta * ra.z;   // theta
}

}  // namespace mte
