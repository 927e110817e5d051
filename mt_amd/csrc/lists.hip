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
#include "rebuild_common.h"

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
    RebuildOwn o{};
    bool skip = true;
    if (active) {
        o = rebuild_own(S.xyz[b], S.fn[b], S.fp1[b], S.fp2[b], T.mon_type[b], T.harm_raw[b], li, b, base,
                        T.tbox[tl.x]);
        skip = T.extra[b] != 0;
    }
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
        if (active && !skip)
            for (int jj = 0; jj < m; jj++)
                rebuild_pair(T, o, j0 + jj, sx[jj], sn[jj], s1[jj], s2[jj], smt[jj], P.ljpairscutoff, do_lj,
                             do_bonded);
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
