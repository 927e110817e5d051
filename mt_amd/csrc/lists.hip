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
// instead of once per (i, j).  Within a tile, j is visited in groups of 8
// consecutive beads with their bounding box: a group whose box lies beyond
// the candidate radius (the LJ list cutoff or the 6.75 nm bonded prefilter,
// whichever is larger) from bead i holds none of i's entries, and a wave
// skips it when all its lanes do.  The surviving j pass a centre-distance
// test into a per-lane candidate buffer, and only the buffered candidates
// (a handful per bead) go through the full tests of rebuild_pair, in
// ascending j, so the lists are those of the plain all-pairs scan.
// Capacity overflows raise T.overflow instead of writing out of bounds (the
// reference does not check, :589-592).
#include "engine_internal.h"
#include "rebuild_common.h"

namespace mte {

constexpr int LIST_TILE = 256;
constexpr int LIST_GROUP = 8;                        // beads per bounding box
constexpr int LIST_NGROUPS = LIST_TILE / LIST_GROUP;
constexpr int LIST_CAND = 24;                        // buffered candidates per lane before a flush

__global__ void __launch_bounds__(LIST_TILE)
    k_rebuild(mte_params P, DevTopo T, DevState S, const int2 *tiles, int do_lj, int do_bonded) {
    __shared__ float4 sx[LIST_TILE], sn[LIST_TILE], s1[LIST_TILE], s2[LIST_TILE];
    __shared__ int smt[LIST_TILE];
    __shared__ float4 glo[LIST_NGROUPS], ghi[LIST_NGROUPS];
    __shared__ uint8_t cand[LIST_CAND * LIST_TILE];   // [slot][lane]: positions within the j tile
    const int2 tl = tiles[blockIdx.x];
    const int base = T.tbase[tl.x], n = T.tsize[tl.x];
    const float box = T.tbox[tl.x];
    const int li = tl.y + threadIdx.x;
    const bool active = li < n;
    const int b = base + li;
    RebuildOwn o{};
    bool skip = true;
    if (active) {
        o = rebuild_own(S.xyz[b], S.fn[b], S.fp1[b], S.fp2[b], T.mon_type[b], T.harm_raw[b], li, b, base, box);
        skip = T.extra[b] != 0;
    }
    const float rc = rebuild_radius(P.ljpairscutoff, do_lj, do_bonded);
    const float rc2 = rc * rc;
    uint8_t *mine = cand + threadIdx.x;
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
        if (threadIdx.x < LIST_NGROUPS && threadIdx.x * LIST_GROUP < m) {
            const int g0 = threadIdx.x * LIST_GROUP, g1 = min(m, g0 + LIST_GROUP);
            float3 lo = f3(sx[g0]), hi = lo;
            for (int q = g0 + 1; q < g1; q++) {
                const float3 x = f3(sx[q]);
                lo = make_float3(fminf(lo.x, x.x), fminf(lo.y, x.y), fminf(lo.z, x.z));
                hi = make_float3(fmaxf(hi.x, x.x), fmaxf(hi.y, x.y), fmaxf(hi.z, x.z));
            }
            glo[threadIdx.x] = make_float4(lo.x, lo.y, lo.z, 0.f);
            ghi[threadIdx.x] = make_float4(hi.x, hi.y, hi.z, 0.f);
        }
        __syncthreads();
        if (active && !skip) {
            int nc = 0;
            auto flush = [&]() {
                for (int q = 0; q < nc; q++) {
                    const int jj = mine[q * LIST_TILE];
                    rebuild_pair(T, o, j0 + jj, sx[jj], sn[jj], s1[jj], s2[jj], smt[jj], P.ljpairscutoff, do_lj,
                                 do_bonded);
                }
                nc = 0;
            };
            for (int g = 0; g * LIST_GROUP < m; g++) {
                // periodic trajectories test every group (the boxes are not minimum-imaged)
                if (box == 0.f && !tile_within(o.ri, glo[g], ghi[g], rc)) continue;
                const int g1 = min(m, (g + 1) * LIST_GROUP);
                for (int jj = g * LIST_GROUP; jj < g1; jj++) {
                    const float3 d = min_image(sub3(o.ri, f3(sx[jj])), box);
                    if (d.x * d.x + d.y * d.y + d.z * d.z < rc2) {
                        mine[nc * LIST_TILE] = (uint8_t)jj;
                        if (++nc == LIST_CAND) flush();
                    }
                }
            }
            flush();
        }
        __syncthreads();
    }
    if (active) rebuild_finish(T, o, do_lj, do_bonded);
}

hipError_t launch_rebuild(mte_engine *e) {
    if (e->use_cells) return launch_rebuild_cells(e);
    const int do_lj = e->P.lj_on ? 1 : 0;
    const int do_bonded = e->P.is_assembly ? 1 : 0;
    if (!do_lj && !do_bonded) return hipSuccess;
    hipLaunchKernelGGL(k_rebuild, dim3(e->ntiles), dim3(LIST_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                       e->tiles, do_lj, do_bonded);
    return hipGetLastError();
}

}  // namespace mte
