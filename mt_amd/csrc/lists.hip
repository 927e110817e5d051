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
// Work layout: one workgroup per (trajectory, 256 rebuild positions); a
// position stands for the bead k_rebuild_order put there when the state was
// last uploaded, which orders each trajectory's beads along a Morton curve of
// cells so that the 64 beads of a wave sit close together and skip far
// neighbour groups together.  The j
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
constexpr int LIST_SUPER = 64;                       // beads per super-group box
constexpr int LIST_CAND = 32;                        // candidate slots per lane
constexpr int TRAJ_THREADS = 1024;                   // k_rebuild_traj workgroup (at most)
constexpr int TRAJ_CAND = 24;                        // its candidate slots per lane
constexpr int ORDER_MAX = REBUILD_ORDER_MAX;         // larger trajectories keep the index order
constexpr int ORDER_THREADS = 1024;

// 7-bit coordinate spread to every third bit (Morton interleave).
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v &= 0x7fu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Rebuild order of each trajectory: beads sorted by the Morton code of their
// cell, ties by index.  Cells are counted from the trajectory's lower
// bounding-box corner, with side `side` or, for a trajectory wider than 127
// such cells, extent / 127, so no coordinate wraps: a wrap would interleave
// beads from opposite ends of the trajectory (on a tube centred on the axis,
// x = -0.1 would fall in cell 127), and 64 consecutive beads would no
// longer be a compact patch.  One workgroup per trajectory; keys
// (code << 11 | bead) bitonic-sorted in LDS.  Any order gives the same
// lists; this one makes the bounding-box skips of k_rebuild coherent across
// a wave and the I tiles of k_tea_pairs_mfma compact.
__global__ void __launch_bounds__(ORDER_THREADS) k_rebuild_order(DevTopo T, DevState S, float side, int *order) {
    __shared__ uint32_t key[ORDER_MAX];
    __shared__ float red[6][ORDER_THREADS / 64];
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    if (n > ORDER_MAX) {
        for (int p = threadIdx.x; p < n; p += blockDim.x) order[base + p] = p;
        return;
    }
    // bounding box: per-lane, then across the wave (xor butterfly), then the waves
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float4 x = S.xyz[base + i];
        b[0] = fminf(b[0], x.x), b[1] = fminf(b[1], x.y), b[2] = fminf(b[2], x.z);
        b[3] = fmaxf(b[3], x.x), b[4] = fmaxf(b[4], x.y), b[5] = fmaxf(b[5], x.z);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const float o = __shfl_xor(b[c], off);
            b[c] = c < 3 ? fminf(b[c], o) : fmaxf(b[c], o);
        }
    const int nw = (int)(blockDim.x + 63) / 64;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < 6; c++) red[c][threadIdx.x >> 6] = b[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 6; c++) {
        b[c] = red[c][0];
        for (int w = 1; w < nw; w++) b[c] = c < 3 ? fminf(b[c], red[c][w]) : fmaxf(b[c], red[c][w]);
    }
    const float ext = fmaxf(fmaxf(b[3] - b[0], b[4] - b[1]), b[5] - b[2]);
    const float inv = 1.f / fmaxf(side, ext * (1.f / 127.f));   // NaN extent: fmaxf keeps side
    int N = 1;
    while (N < n) N <<= 1;
    // cell of one coordinate in [0, 127]; a NaN coordinate goes to cell 0
    auto cell = [&](float v, float lo) { return (uint32_t)fminf(fmaxf((v - lo) * inv, 0.f), 127.f); };
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        uint32_t k = 0xffffffffu;   // padding sorts last
        if (i < n) {
            const float4 x = S.xyz[base + i];
            const uint32_t cx = cell(x.x, b[0]), cy = cell(x.y, b[1]), cz = cell(x.z, b[2]);
            k = ((spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2)) << 11) | (uint32_t)i;
        }
        key[i] = k;
    }
    __syncthreads();
    for (int k = 2; k <= N; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = threadIdx.x; q < N / 2; q += blockDim.x) {
                const int a = 2 * q - (q & (j - 1)), c = a + j;   // a has bit j clear
                const uint32_t ka = key[a], kc = key[c];
                if ((ka > kc) == ((a & k) == 0)) {
                    key[a] = kc;
                    key[c] = ka;
                }
            }
            __syncthreads();
        }
    for (int p = threadIdx.x; p < n; p += blockDim.x) order[base + p] = (int)(key[p] & (ORDER_MAX - 1));
}

// Bounding boxes of the m beads in LDS: groups of LIST_GROUP consecutive
// beads, and super-groups of LIST_SUPER (the union of their groups' boxes).
// Ends with a barrier.
__device__ __forceinline__ void group_boxes(const float4 *sx, int m, float4 *glo, float4 *ghi, float4 *slo,
                                            float4 *shi) {
    const int ng = (m + LIST_GROUP - 1) / LIST_GROUP, nsg = (m + LIST_SUPER - 1) / LIST_SUPER;
    for (int g = threadIdx.x; g < ng; g += blockDim.x) {
        const int g0 = g * LIST_GROUP, g1 = min(m, g0 + LIST_GROUP);
        float3 lo = f3(sx[g0]), hi = lo;
        for (int q = g0 + 1; q < g1; q++) {
            const float3 x = f3(sx[q]);
            lo = make_float3(fminf(lo.x, x.x), fminf(lo.y, x.y), fminf(lo.z, x.z));
            hi = make_float3(fmaxf(hi.x, x.x), fmaxf(hi.y, x.y), fmaxf(hi.z, x.z));
        }
        glo[g] = make_float4(lo.x, lo.y, lo.z, 0.f);
        ghi[g] = make_float4(hi.x, hi.y, hi.z, 0.f);
    }
    __syncthreads();
    constexpr int GPS = LIST_SUPER / LIST_GROUP;
    for (int sg = threadIdx.x; sg < nsg; sg += blockDim.x) {
        float4 lo = glo[sg * GPS], hi = ghi[sg * GPS];
        for (int g = sg * GPS + 1; g < min(ng, (sg + 1) * GPS); g++) {
            const float4 l = glo[g], h = ghi[g];
            lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), 0.f);
            hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), 0.f);
        }
        slo[sg] = lo;
        shi[sg] = hi;
    }
    __syncthreads();
}

// Candidates of the bead at ri among the m beads in LDS: the beads of every
// group (inside a super-group) whose box comes within rc and whose centre
// distance is below rc, appended in ascending order to the lane's buffer
// mine[slot * stride] (each slot is written and kept only if the test
// passes: no branch per bead); the buffer is flushed through the full tests
// whenever another group might not fit in its `cap` slots.
template <bool PERIODIC, class CandT, class Flush>
__device__ __forceinline__ void scan_groups(const float4 *sx, const float4 *glo, const float4 *ghi, const float4 *slo,
                                            const float4 *shi, int m, float3 ri, float box, float rc, float rc2,
                                            CandT *mine, int stride, int cap, int &nc, const Flush &flush) {
    constexpr int GPS = LIST_SUPER / LIST_GROUP;
    const int ng = (m + LIST_GROUP - 1) / LIST_GROUP;
    for (int sg = 0; sg * LIST_SUPER < m; sg++) {
        if (!PERIODIC && !tile_within(ri, slo[sg], shi[sg], rc)) continue;
        const int gend = min(ng, (sg + 1) * GPS);
        for (int g = sg * GPS; g < gend; g++) {
            if (!PERIODIC && !tile_within(ri, glo[g], ghi[g], rc)) continue;
            const int g0 = g * LIST_GROUP;
            auto test = [&](int jj) {
                float3 d = sub3(ri, f3(sx[jj]));
                if (PERIODIC) d = min_image(d, box);
                mine[nc * stride] = (CandT)jj;
                nc += (d.x * d.x + d.y * d.y + d.z * d.z < rc2) ? 1 : 0;
            };
            if (g0 + LIST_GROUP <= m) {
#pragma unroll
                for (int u = 0; u < LIST_GROUP; u++) test(g0 + u);
            } else {
                for (int jj = g0; jj < m; jj++) test(jj);
            }
            if (nc > cap - LIST_GROUP) flush();
        }
    }
}

__global__ void __launch_bounds__(LIST_TILE)
    k_rebuild(mte_params P, DevTopo T, DevState S, const int2 *tiles, const int *order, int do_lj, int do_bonded) {
    __shared__ float4 sx[LIST_TILE], sn[LIST_TILE], s1[LIST_TILE], s2[LIST_TILE];
    __shared__ int smt[LIST_TILE];
    __shared__ float4 glo[LIST_NGROUPS], ghi[LIST_NGROUPS], slo[LIST_TILE / LIST_SUPER], shi[LIST_TILE / LIST_SUPER];
    __shared__ uint8_t cand[LIST_CAND * LIST_TILE];   // [slot][lane]: positions within the j tile
    const int2 tl = tiles[blockIdx.x];
    const int base = T.tbase[tl.x], n = T.tsize[tl.x];
    const float box = T.tbox[tl.x];
    const int pos = tl.y + threadIdx.x;
    const bool active = pos < n;
    const int li = active ? order[base + pos] : 0;
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
        group_boxes(sx, m, glo, ghi, slo, shi);
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
            // periodic trajectories test every group (the boxes are not minimum-imaged)
            if (box > 0.f)
                scan_groups<true>(sx, glo, ghi, slo, shi, m, o.ri, box, rc, rc2, mine, LIST_TILE, LIST_CAND, nc, flush);
            else
                scan_groups<false>(sx, glo, ghi, slo, shi, m, o.ri, box, rc, rc2, mine, LIST_TILE, LIST_CAND, nc, flush);
            flush();
        }
        __syncthreads();
    }
    if (active) rebuild_finish(T, o, do_lj, do_bonded);
}

// Candidates of the bead at ri among the m beads of a trajectory held in LDS
// in rebuild order (sm[pos].w: the bead's local index): every bead of a
// group (inside a super-group) whose box comes within rc and whose centre
// distance is below rc.  They come out in rebuild order, not by index, so
// nothing is flushed during the scan: nc counts them all, the first `cap`
// are kept in mine[slot * stride], and every later write lands in slot
// `cap`, which the buffer holds beyond them.
template <bool PERIODIC>
__device__ __forceinline__ void scan_cands(const float4 *sm, const float4 *glo, const float4 *ghi, const float4 *slo,
                                           const float4 *shi, int m, float3 ri, float box, float rc, float rc2,
                                           uint16_t *mine, int stride, int cap, int &nc) {
    constexpr int GPS = LIST_SUPER / LIST_GROUP;
    const int ng = (m + LIST_GROUP - 1) / LIST_GROUP;
    for (int sg = 0; sg * LIST_SUPER < m; sg++) {
        if (!PERIODIC && !tile_within(ri, slo[sg], shi[sg], rc)) continue;
        const int gend = min(ng, (sg + 1) * GPS);
        for (int g = sg * GPS; g < gend; g++) {
            if (!PERIODIC && !tile_within(ri, glo[g], ghi[g], rc)) continue;
            const int g0 = g * LIST_GROUP;
            auto test = [&](int jj) {
                const float4 q = sm[jj];
                float3 d = sub3(ri, f3(q));
                if (PERIODIC) d = min_image(d, box);
                mine[min(nc, cap) * stride] = (uint16_t)__float_as_int(q.w);
                nc += (d.x * d.x + d.y * d.y + d.z * d.z < rc2) ? 1 : 0;
            };
            if (g0 + LIST_GROUP <= m) {
#pragma unroll
                for (int u = 0; u < LIST_GROUP; u++) test(g0 + u);
            } else {
                for (int jj = g0; jj < m; jj++) test(jj);
            }
        }
    }
}

// The lane's nc <= TRAJ_CAND candidates in ascending order: a bitonic
// network over SORT_W registers (padding sorts last), with no dependent LDS
// accesses.
constexpr int SORT_W = 32;
static_assert(TRAJ_CAND <= SORT_W, "the sorting network covers the candidate buffer");
__device__ __forceinline__ void sort_cands(uint16_t *mine, int stride, int nc) {
    uint32_t v[SORT_W];
#pragma unroll
    for (int q = 0; q < SORT_W; q++) v[q] = (q < TRAJ_CAND && q < nc) ? mine[q * stride] : 0xffffu;
#pragma unroll
    for (int k = 2; k <= SORT_W; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < SORT_W; i++) {
                const int l = i ^ j;
                if (l > i) {
                    const uint32_t a = min(v[i], v[l]), b = max(v[i], v[l]);
                    const bool up = (i & k) == 0;
                    v[i] = up ? a : b;
                    v[l] = up ? b : a;
                }
            }
#pragma unroll
    for (int q = 0; q < TRAJ_CAND; q++)
        if (q < nc) mine[q * stride] = (uint16_t)v[q];
}

// The same rebuild with one workgroup per trajectory and the whole
// trajectory in LDS (positions, rotated offsets, monomer types, a uint16
// candidate buffer per lane), used when the largest trajectory fits
// (rebuild_traj_plan).  Every bead's data crosses from L2 once instead of
// once per 256-bead i tile, and the neighbour search runs over a second copy
// of the positions in rebuild order, whose groups of 8 are compact patches
// (index order runs along a protofilament, so its groups are 28 nm rods
// that pass the box test for most of a wave's lanes).  Each lane sorts its
// candidates by index before the full tests, so the lists come out in the
// reference's order; a bead with more than TRAJ_CAND candidates tests every
// bead of its trajectory in index order instead.
__global__ void __launch_bounds__(TRAJ_THREADS)
    k_rebuild_traj(mte_params P, DevTopo T, DevState S, const int *order, int do_lj, int do_bonded) {
    extern __shared__ float4 lds[];
    const int t = blockIdx.x;
    const int base = T.tbase[t], n = T.tsize[t];
    const int ng = (n + LIST_GROUP - 1) / LIST_GROUP, nsg = (n + LIST_SUPER - 1) / LIST_SUPER;
    float4 *sx = lds, *sm = sx + n, *sn = sm + n, *s1 = sn + n, *s2 = s1 + n;
    float4 *glo = s2 + n, *ghi = glo + ng, *slo = ghi + ng, *shi = slo + nsg;
    int *smt = reinterpret_cast<int *>(shi + nsg);
    uint16_t *cand = reinterpret_cast<uint16_t *>(smt + n);   // [slot][thread], TRAJ_CAND + 1 slots
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        const int j = order[base + q];
        const float4 xj = S.xyz[base + j];
        sm[q] = make_float4(xj.x, xj.y, xj.z, __int_as_float(j));
        sx[q] = S.xyz[base + q];
        sn[q] = S.fn[base + q];
        s1[q] = S.fp1[base + q];
        s2[q] = S.fp2[base + q];
        smt[q] = T.mon_type[base + q];
    }
    __syncthreads();
    group_boxes(sm, n, glo, ghi, slo, shi);
    const float box = T.tbox[t];
    const float rc = rebuild_radius(P.ljpairscutoff, do_lj, do_bonded);
    const float rc2 = rc * rc;
    uint16_t *mine = cand + threadIdx.x;
    const int stride = (int)blockDim.x;
    for (int pos = threadIdx.x; pos < n; pos += blockDim.x) {
        const int li = __float_as_int(sm[pos].w);
        const int b = base + li;
        RebuildOwn o = rebuild_own(sx[li], sn[li], s1[li], s2[li], smt[li], T.harm_raw[b], li, b, base, box);
        if (!T.extra[b]) {
            int nc = 0;
            if (box > 0.f)
                scan_cands<true>(sm, glo, ghi, slo, shi, n, o.ri, box, rc, rc2, mine, stride, TRAJ_CAND, nc);
            else
                scan_cands<false>(sm, glo, ghi, slo, shi, n, o.ri, box, rc, rc2, mine, stride, TRAJ_CAND, nc);
            // the full tests; LJ-only rebuilds read no rotated offsets or types
            auto pair = [&](int jj) {
                if (do_bonded)
                    rebuild_pair(T, o, jj, sx[jj], sn[jj], s1[jj], s2[jj], smt[jj], P.ljpairscutoff, do_lj, true);
                else
                    rebuild_pair(T, o, jj, sx[jj], float4{}, float4{}, float4{}, 0, P.ljpairscutoff, do_lj, false);
            };
            if (nc <= TRAJ_CAND) {
                sort_cands(mine, stride, nc);
                for (int q = 0; q < nc; q++) pair(mine[q * stride]);
            } else {
                for (int jj = 0; jj < n; jj++) {
                    float3 d = sub3(o.ri, f3(sx[jj]));
                    if (box > 0.f) d = min_image(d, box);
                    if (d.x * d.x + d.y * d.y + d.z * d.z < rc2) pair(jj);
                }
            }
        }
        rebuild_finish(T, o, do_lj, do_bonded);
    }
}

// Workgroup size and dynamic LDS of k_rebuild_traj for trajectories of up to
// nmax beads; false when they do not fit (the tiled kernel is used).
bool rebuild_traj_plan(int nmax, int lds_limit, int *threads, size_t *lds) {
    if (nmax <= 0 || nmax > 65535) return false;
    const int th = std::min(TRAJ_THREADS, ((nmax + 63) / 64) * 64);
    const size_t ng = (nmax + LIST_GROUP - 1) / LIST_GROUP, nsg = (nmax + LIST_SUPER - 1) / LIST_SUPER;
    const size_t bytes = sizeof(float4) * (5 * (size_t)nmax + 2 * ng + 2 * nsg) + sizeof(int) * (size_t)nmax +
                         sizeof(uint16_t) * (size_t)(TRAJ_CAND + 1) * th;
    if (bytes > (size_t)lds_limit) return false;
    *threads = th;
    *lds = (bytes + 15) & ~(size_t)15;
    return true;
}

hipError_t launch_rebuild(mte_engine *e) {
    if (e->use_cells) return launch_rebuild_cells(e);
    const int do_lj = e->P.lj_on ? 1 : 0;
    const int do_bonded = e->P.is_assembly ? 1 : 0;
    if (!do_lj && !do_bonded) return hipSuccess;
    if (e->rb_traj_threads > 0) {
        const hipError_t s = hipFuncSetAttribute((const void *)k_rebuild_traj,
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)e->rb_traj_lds);
        if (s != hipSuccess) return s;
        hipLaunchKernelGGL(k_rebuild_traj, dim3(e->ntraj), dim3(e->rb_traj_threads), e->rb_traj_lds, e->stream, e->P,
                           e->T, e->S[e->cur], e->rb_order, do_lj, do_bonded);
    } else {
        hipLaunchKernelGGL(k_rebuild, dim3(e->ntiles), dim3(LIST_TILE), 0, e->stream, e->P, e->T, e->S[e->cur],
                           e->tiles, e->rb_order, do_lj, do_bonded);
    }
    return hipGetLastError();
}

// Spatial rebuild order from the current state.  Run when a state is set, not
// per rebuild: beads move little between uploads, and a stale order costs
// only skips, never entries.
hipError_t launch_rebuild_order(mte_engine *e) {
    const float side = rebuild_radius(e->P.ljpairscutoff, true, true);
    hipLaunchKernelGGL(k_rebuild_order, dim3(e->ntraj), dim3(ORDER_THREADS), 0, e->stream, e->T, e->S[e->cur], side,
                       e->rb_order);
    return hipGetLastError();
}

}  // namespace mte
