// rebuild_common.h -- the per-(i, j) candidate tests of a list rebuild,
// shared by the rebuild kernels of lists.hip (j tiles in LDS, or the whole
// trajectory in LDS) and the cell-list rebuild of cells.hip.
//
// Restates pairs_kernel (src/compute_cuda.cu:527-674) and LJ_kernel
// (:913-940); see lists.hip for the semantics.
#pragma once
#include "device_common.h"

namespace mte {

// Radius beyond which no pair enters any list: the LJ list cutoff and the
// bonded prefilter, widened so float rounding of a distance cannot cross it.
__host__ __device__ inline float rebuild_radius(float ljcut, bool do_lj, bool do_bonded) {
    float r = 0.f;
    if (do_lj) r = fmaxf(r, ljcut);
    if (do_bonded) r = fmaxf(r, BOND_PREFILTER);
    return r * 1.0001f + 1e-4f;
}

// sqrtf(d2) < cut, the list test of the reference, with the square root
// evaluated only inside a relative 2^-20 band around cut: outside it the
// comparison of d2 with the band's squared ends decides the same way,
// sqrtf being correctly rounded and monotonic.
__device__ __forceinline__ bool below_cut(float d2, float cut) {
    const float lo = cut * (1.f - 0x1p-20f), hi = cut * (1.f + 0x1p-20f);
    if (d2 < lo * lo) return true;
    if (d2 >= hi * hi) return false;
    return sqrtf(d2) < cut;
}

// True when the box [lo, hi] comes within r of point p (open boundaries).
__device__ __forceinline__ bool tile_within(float3 p, float4 lo, float4 hi, float r) {
    const float ex = fmaxf(fmaxf(lo.x - p.x, p.x - hi.x), 0.f);
    const float ey = fmaxf(fmaxf(lo.y - p.y, p.y - hi.y), 0.f);
    const float ez = fmaxf(fmaxf(lo.z - p.z, p.z - hi.z), 0.f);
    return ex * ex + ey * ey + ez * ez < r * r;
}

struct RebuildOwn {
    float3 ri, ani, a1, a2;   // position, R_MON n_i, R p1, R p2
    float rp;                  // R_MON used by pairs_kernel
    float box;                 // periodic side (0: open)
    float ibox;                // inv_side(box)
    int mti, hraw, hpart, li, b, base;
    int nlon, nlat, nlj, ovf;
};

__device__ __forceinline__ RebuildOwn rebuild_own(float4 x, float4 n, float4 p1, float4 p2, int mti, int hraw, int li,
                                                  int b, int base, float box) {
    RebuildOwn o;
    o.box = box;
    o.ibox = inv_side(box);
    o.ri = f3(x);
    // pairs_kernel takes R_MON and the excluded partner from the first
    // harmonic slot, including the -1 fill of partnerless beads (:548-551, :565).
    o.rp = (hraw < 0) ? -RMON : RMON;
    o.ani = make_float3(o.rp * n.x, o.rp * n.y, o.rp * n.z);
    o.a1 = f3(p1);
    o.a2 = f3(p2);
    o.mti = mti;
    o.hraw = hraw;
    o.hpart = abs(hraw);
    o.li = li;
    o.b = b;
    o.base = base;
    o.nlon = o.nlat = o.nlj = o.ovf = 0;
    return o;
}

// Test local bead j against bead o; append to the k-major global lists.
__device__ __forceinline__ void rebuild_pair(const DevTopo &T, RebuildOwn &o, int j, float4 xj, float4 nj,
                                             float4 p1j, float4 p2j, int mtj, float ljcut, bool do_lj,
                                             bool do_bonded) {
    const int nb = T.nb;
    const float3 dij = min_image_pre(sub3(o.ri, f3(xj)), o.box, o.ibox);
    const float d2 = dij.x * dij.x + dij.y * dij.y + dij.z * dij.z;
    if (do_lj && below_cut(d2, ljcut) && o.li != j) {
        if (o.nlj < T.max_lj) T.lj[o.nlj * nb + o.b] = o.base + j;
        else o.ovf = 1;
        o.nlj++;
    }
    if (!do_bonded || d2 >= BOND_PREFILTER * BOND_PREFILTER || o.hpart == j) return;
    const float3 z3 = make_float3(0, 0, 0);
    const float2 z2 = make_float2(0, 0);
    if (o.mti != mtj) {
        const Grad G = gap_grad(dij, o.ani, make_float3(-o.rp * nj.x, -o.rp * nj.y, -o.rp * nj.z), z3, z2);
        if (G.dr < PAIR_CUT) {
            // stored (h < 0 ? j : -j) and decoded "negative -> +r_mon": R_MON > 0
            // iff h >= 0 and j != 0 (-0 == 0 decodes as -r_mon).
            const int rpos = (o.hraw >= 0 && j != 0) ? 1 : 0;
            if (o.nlon < MAXLON) T.lon[o.nlon * nb + o.b] = ((o.base + j) << 1) | rpos;
            else o.ovf = 1;
            o.nlon++;
        }
    }
    if (o.li != j) {
        // orientation 0: i uses p1, j uses p2 (stored negative)
        Grad G = gap_grad(dij, o.a1, f3(p2j), z3, z2);
        if (G.dr < PAIR_CUT) {
            if (o.nlat < MAXLAT) T.lat[o.nlat * nb + o.b] = ((o.base + j) << 1) | 1;
            else o.ovf = 1;
            o.nlat++;
        }
        // orientation 1: i uses p2, j uses p1 (stored positive)
        G = gap_grad(dij, o.a2, f3(p1j), z3, z2);
        if (G.dr < PAIR_CUT) {
            if (o.nlat < MAXLAT) T.lat[o.nlat * nb + o.b] = ((o.base + j) << 1);
            else o.ovf = 1;
            o.nlat++;
        }
    }
}

__device__ __forceinline__ void rebuild_finish(const DevTopo &T, const RebuildOwn &o, bool do_lj, bool do_bonded) {
    if (do_lj) T.lj_n[o.b] = min(o.nlj, T.max_lj);
    if (do_bonded) {
        T.lon_n[o.b] = min(o.nlon, MAXLON);
        T.lat_n[o.b] = min(o.nlat, MAXLAT);
    }
    if (o.ovf) atomicOr(T.overflow, 1);
}

}  // namespace mte
