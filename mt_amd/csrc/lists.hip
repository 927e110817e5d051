// lists.hip -- neighbour-list rebuild (streamed mode).
//
// Restates pairs_kernel (src/compute_cuda.cu:527-674) and LJ_kernel
// (:913-940): for bead i, every j of its trajectory in increasing j is tested
// for (a) LJ: centre distance < LJPairsCutoff, j != i (extra beads included,
// as the reference does); (b) longitudinal: opposite mon_type, j not the
// dimer partner, axial attachment gap < 2.5 nm; (c) lateral: j != i, j not
// the partner, either orientation's lateral gap < 2.5 nm.  Entries come out
// in the reference's order (ascending j; lateral orientation 0 before 1).
//
// Work layout: one workgroup per (trajectory, 256-bead tile of i); the j
// range of the trajectory streams through LDS in 256-bead tiles, so each
// neighbour's position and rotated offsets are read from HBM once per tile
// instead of once per (i, j).  A centre-distance prefilter (no bonded
// candidate can lie beyond 6.75 nm, device_common.h) skips the attachment
// gaps for far pairs.  Capacity overflows raise T.overflow instead of
// writing out of bounds (the reference does not check, :589-592).
#include "engine_internal.h"

namespace mte {

constexpr int LIST_TILE = 256;

__global__ void __launch_bounds__(LIST_TILE)
    k_rebuild(mte_params P, DevTopo T, DevState S, const int2 *tiles, int do_lj, int do_bonded) {
    __shared__ float4 sx[LIST_TILE], sn[LIST_TILE], s1[LIST_TILE], s2[LIST_TILE];
    __shared__ int smt[LIST_TILE];
    const int2 tl = tiles[blockIdx.x];
    const int base = T.tbase[tl.x], n = T.tsize[tl.x];
    const int li = tl.y + threadIdx.x;
    const bool active = li < n;
    const int b = base + li;
    const int nb = T.nb;

    float3 ri = make_float3(0, 0, 0), ni = ri, a1 = ri, a2 = ri;
    int mti = 0, hraw = -1;
    bool skip = true;
    if (active) {
        ri = f3(S.xyz[b]);
        ni = f3(S.fn[b]);
        a1 = f3(S.fp1[b]);
        a2 = f3(S.fp2[b]);
        mti = T.mon_type[b];
        hraw = T.harm_raw[b];
        skip = T.extra[b] != 0;
    }
    // pairs_kernel takes R_MON and the excluded partner from the first
    // harmonic slot, including the -1 fill of partnerless beads (:548-551, :565).
    const float rp = (hraw < 0) ? -RMON : RMON;
    const int hpart = abs(hraw);
    const float3 ani = make_float3(rp * ni.x, rp * ni.y, rp * ni.z);
    const float cut2 = P.ljpairscutoff;
    const float pre2 = BOND_PREFILTER * BOND_PREFILTER;
    const float3 z3 = make_float3(0, 0, 0);
    const float2 z2 = make_float2(0, 0);
    int nlon = 0, nlat = 0, nlj = 0, ovf = 0;

    for (int j0 = 0; j0 < n; j0 += LIST_TILE) {
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            sx[threadIdx.x] = S.xyz[base + jl];
            sn[threadIdx.x] = S.fn[base + jl];
            s1[threadIdx.x] = S.fp1[base + jl];
            s2[threadIdx.x] = S.fp2[base + jl];
            smt[threadIdx.x] = T.mon_type[base + jl];
        }
        __syncthreads();
        const int m = min(LIST_TILE, n - j0);
        if (active && !skip) {
            for (int jj = 0; jj < m; jj++) {
                const int j = j0 + jj;
                const float3 rj = f3(sx[jj]);
                const float dx = ri.x - rj.x, dy = ri.y - rj.y, dz = ri.z - rj.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                if (do_lj && sqrtf(d2) < cut2 && li != j) {
                    if (nlj < T.max_lj) T.lj[nlj * nb + b] = base + j;
                    else ovf = 1;
                    nlj++;
                }
                if (!do_bonded || d2 >= pre2 || hpart == j) continue;
                if (mti != smt[jj]) {
                    const float3 nj = f3(sn[jj]);
                    const Grad G = gap_grad(ri, ani, rj, make_float3(-rp * nj.x, -rp * nj.y, -rp * nj.z), z3, z2);
                    if (G.dr < PAIR_CUT) {
                        // stored (h < 0 ? j : -j) and decoded "negative -> +r_mon":
                        // R_MON > 0 iff h >= 0 and j != 0 (-0 == 0 decodes as -r_mon).
                        const int rpos = (hraw >= 0 && j != 0) ? 1 : 0;
                        if (nlon < MAXLON) T.lon[nlon * nb + b] = ((base + j) << 1) | rpos;
                        else ovf = 1;
                        nlon++;
                    }
                }
                if (li != j) {
                    // orientation 0: i uses p1, j uses p2 (stored negative)
                    Grad G = gap_grad(ri, a1, rj, f3(s2[jj]), z3, z2);
                    if (G.dr < PAIR_CUT) {
                        if (nlat < MAXLAT) T.lat[nlat * nb + b] = ((base + j) << 1) | 1;
                        else ovf = 1;
                        nlat++;
                    }
                    // orientation 1: i uses p2, j uses p1 (stored positive)
                    G = gap_grad(ri, a2, rj, f3(s1[jj]), z3, z2);
                    if (G.dr < PAIR_CUT) {
                        if (nlat < MAXLAT) T.lat[nlat * nb + b] = ((base + j) << 1);
                        else ovf = 1;
                        nlat++;
                    }
                }
            }
        }
        __syncthreads();
    }
    if (active) {
        if (do_lj) T.lj_n[b] = min(nlj, T.max_lj);
        if (do_bonded) {
            T.lon_n[b] = min(nlon, MAXLON);
            T.lat_n[b] = min(nlat, MAXLAT);
        }
        if (ovf) atomicOr(T.overflow, 1);
    }
}

hipError_t launch_rebuild(mte_engine *e) {
    const int do_lj = e->P.lj_on ? 1 : 0;
    const int do_bonded = e->P.is_assembly ? 1 : 0;
    if (!do_lj && !do_bonded) return hipSuccess;
    hipLaunchKernelGGL(k_rebuild, dim3(e->ntiles), dim3(LIST_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                       e->tiles, do_lj, do_bonded);
    return hipGetLastError();
}

}  // namespace mte
