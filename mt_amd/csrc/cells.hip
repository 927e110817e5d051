// cells.hip -- cell-list neighbour search (sort and scan) for list rebuilds
// of large trajectories.
//
// The reference rebuilds its lists by scanning every bead of the trajectory
// for every bead (pairs_kernel, src/compute_cuda.cu:527-674; LJ_kernel,
// :913-940), O(N^2) per trajectory.  Here the rebuild is:
//   1. assign: cell coordinates floor(x / cs) per bead (wrapped in periodic
//      boxes), hashed into a per-trajectory table of 2^k buckets; count;
//   2. scan: one exclusive scan over the buckets of all trajectories (whose
//      tables and beads are both laid out trajectory after trajectory)
//      turns bucket counts into first slots;
//   3. scatter: beads into bucket order (counting sort), with their
//      positions copied into that order for the searches;
//   4. search: each bead visits the 27 neighbouring cells, keeps the beads of
//      exactly those cells within the candidate radius, sorts the candidates
//      by index and runs the reference's per-pair tests on them in that
//      order, so the lists come out identical to the all-pairs scan (same
//      entries, same order).
// A hash collision only adds beads that the exact cell check then rejects.
#include "engine_internal.h"
#include "rebuild_common.h"

namespace mte {

constexpr int CELL_BLOCK = 64;
constexpr int CAND_CAP = 128;   // candidates per bead within the candidate radius

__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, int cz) {
    return ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
}

struct CellGeom {
    float cs;      // cell side
    int dims;      // cells per side in a periodic box (0: open)
};

// Cells of side >= `side` (searched K cells either way, so side = r / K covers
// a radius r); a periodic box has at least 2K + 1 cells per side, in which
// case the stencil spans the whole box.
__device__ __forceinline__ CellGeom cell_geom(float box, float side, int K) {
    CellGeom g;
    if (box > 0.f) {
        g.dims = max(2 * K + 1, (int)floorf(box / side));
        g.cs = box / (float)g.dims;
    } else {
        g.dims = 0;
        g.cs = side;
    }
    return g;
}

__device__ __forceinline__ int3 cell_of(float3 r, float box, const CellGeom &g) {
    if (g.dims > 0) {
        const float wx = r.x - box * floorf(r.x / box), wy = r.y - box * floorf(r.y / box),
                    wz = r.z - box * floorf(r.z / box);
        return make_int3(min(g.dims - 1, (int)(wx / g.cs)), min(g.dims - 1, (int)(wy / g.cs)),
                         min(g.dims - 1, (int)(wz / g.cs)));
    }
    return make_int3((int)floorf(r.x / g.cs), (int)floorf(r.y / g.cs), (int)floorf(r.z / g.cs));
}

// The bead's slot within its bucket comes from one atomic per bucket and
// wave: the wave's lanes are grouped by bucket (one round per distinct
// bucket, scalar work), each group's first lane adds the group's size, and
// the lanes of a group take consecutive slots from the returned count.
// Neighbouring beads of a microtubule share cells, so most of a wave's
// lanes fall in a few buckets.  The order of the beads within a bucket is
// not fixed (nor was it with one atomic per bead); every user of the tables
// orders what it finds there by bead index or by distance.
__global__ void __launch_bounds__(256)
    k_cell_assign(DevTopo T, DevState S, CellTables C, float side, int K) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    const bool valid = b < T.nb;
    int bucket = -1;
    if (valid) {
        const int t = T.tid[b];
        const float box = T.tbox[t];
        const CellGeom g = cell_geom(box, side, K);
        const int3 c = cell_of(f3(S.xyz[b]), box, g);
        bucket = C.toff[t] + (int)(cell_hash(c.x, c.y, c.z) & (uint32_t)(C.tsize[t] - 1));
        C.cell[b] = make_int4(c.x, c.y, c.z, bucket);
    }
    const int lane = threadIdx.x & 63;
    uint64_t todo = __ballot(valid), group = 0;
    int leader = lane;
    while (todo) {
        const int l0 = __ffsll((unsigned long long)todo) - 1;
        const int kb = __builtin_amdgcn_readlane(bucket, l0);
        const uint64_t m = __ballot(valid && bucket == kb);
        if (valid && bucket == kb) {
            leader = l0;
            group = m;
        }
        todo &= ~m;
    }
    int first = 0;
    if (valid && lane == leader) first = atomicAdd(&C.count[bucket], __popcll(group));
    first = __shfl(first, leader);
    if (valid) C.rank[b] = first + __popcll(group & ((1ull << lane) - 1ull));
}

// Exclusive scan of the bucket counts over ALL trajectories at once.  Each
// trajectory's buckets follow the previous trajectory's, and its beads follow
// the previous trajectory's beads, so the global prefix of bucket k is
// directly the first slot of that bucket (tbase[t] + the local prefix).
// Three passes: per-block sums, a scan of the block sums, the in-block scan.
constexpr int SCAN_T = 256, SCAN_PER = CELL_SCAN_SPAN / SCAN_T;

__global__ void __launch_bounds__(SCAN_T) k_scan_partial(const int *count, int n, int *bsum) {
    __shared__ int red[SCAN_T];
    const int g0 = blockIdx.x * CELL_SCAN_SPAN + threadIdx.x;   // lanes read consecutive buckets
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++)
        if (g0 + k * SCAN_T < n) s += count[g0 + k * SCAN_T];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int d = SCAN_T / 2; d > 0; d >>= 1) {
        if (threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// In-place exclusive scan of the nblk block sums (one workgroup).
__global__ void __launch_bounds__(1024) k_scan_top(int *bsum, int nblk) {
    __shared__ int part[1024];
    int carry = 0;
    for (int c0 = 0; c0 < nblk; c0 += 1024) {
        const int i = c0 + threadIdx.x;
        const int v = i < nblk ? bsum[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {
            const int u = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
            __syncthreads();
            part[threadIdx.x] += u;
            __syncthreads();
        }
        if (i < nblk) bsum[i] = carry + part[threadIdx.x] - v;
        const int tot = part[1023];
        __syncthreads();
        carry += tot;
    }
}

// Packed (first slot, count) ranges of the buckets (for the scatter and the
// searches).  The span's counts pass through LDS
// so that global loads and stores are coalesced (lane-consecutive buckets)
// while each thread scans a contiguous run of SCAN_PER; one pad word per 16
// keeps a run's reads off shared banks.
__device__ __forceinline__ int scan_pad(int i) { return i + (i >> 4); }
__global__ void __launch_bounds__(SCAN_T) k_scan_apply(CellTables C) {
    __shared__ int part[SCAN_T];
    __shared__ int cnt[CELL_SCAN_SPAN + CELL_SCAN_SPAN / 16], st[CELL_SCAN_SPAN + CELL_SCAN_SPAN / 16];
    const int n = C.total;
    const int g0 = blockIdx.x * CELL_SCAN_SPAN;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        const int idx = k * SCAN_T + threadIdx.x;
        cnt[scan_pad(idx)] = g0 + idx < n ? C.count[g0 + idx] : 0;
        if (g0 + idx < n) C.count[g0 + idx] = 0;   // ready for the next build's counting atomics
    }
    __syncthreads();
    const int run = threadIdx.x * SCAN_PER;
    int v[SCAN_PER];
    int s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        v[k] = cnt[scan_pad(run + k)];
        s += v[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int d = 1; d < SCAN_T; d <<= 1) {
        const int u = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += u;
        __syncthreads();
    }
    int acc = C.bsum[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        st[scan_pad(run + k)] = acc;
        acc += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_PER; k++) {
        const int idx = k * SCAN_T + threadIdx.x;
        if (g0 + idx < n) C.range[g0 + idx] = make_int2(st[scan_pad(idx)], cnt[scan_pad(idx)]);
    }
}

// Counting-sort scatter; also the bucket-ordered positions the searches
// stream through (w = bead index, bit 31 set for extra beads).
__global__ void __launch_bounds__(256) k_cell_scatter(DevTopo T, DevState S, CellTables C) {
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= T.nb) return;
    const int slot = C.range[C.cell[b].w].x + C.rank[b];
    C.sorted[slot] = b;
    const float4 x = S.xyz[b];
    C.spos[slot] = make_float4(x.x, x.y, x.z, __int_as_float(b | (T.extra[b] ? (int)0x80000000 : 0)));
}

// Candidates of bead b: beads of the (2K+1)^3 neighbouring cells (of side
// >= radius / K) within `radius` (minimum image), as local indices in
// ascending order.  Returns the count, or -1 when more than CAND_CAP were
// found.
__device__ int gather_candidates(const DevTopo &T, const DevState &S, const CellTables &C, int b, float radius,
                                 int K, int *cand) {
    const int t = T.tid[b], base = T.tbase[t];
    const float box = T.tbox[t], ibox = inv_side(box);
    const CellGeom g = cell_geom(box, radius / K, K);
    const int4 me = C.cell[b];
    const float3 ri = f3(S.xyz[b]);
    const float r2max = radius * radius;
    const uint32_t mask = (uint32_t)(C.tsize[t] - 1);
    int n = 0;
    for (int dz = -K; dz <= K; dz++)
        for (int dy = -K; dy <= K; dy++)
            for (int dx = -K; dx <= K; dx++) {
                int cx = me.x + dx, cy = me.y + dy, cz = me.z + dz;
                if (g.dims > 0) {
                    cx = wrap_cell(cx, g.dims);
                    cy = wrap_cell(cy, g.dims);
                    cz = wrap_cell(cz, g.dims);
                }
                const int bucket = C.toff[t] + (int)(cell_hash(cx, cy, cz) & mask);
                const int2 rg = C.range[bucket];
                const int s0 = rg.x, s1 = rg.x + rg.y;
                for (int q = s0; q < s1; q++) {
                    const int j = C.sorted[q];
                    if (j == b) continue;
                    const int4 cj = C.cell[j];
                    if (cj.x != cx || cj.y != cy || cj.z != cz) continue;
                    const float3 d = min_image_pre(sub3(ri, f3(S.xyz[j])), box, ibox);
                    if (d.x * d.x + d.y * d.y + d.z * d.z >= r2max) continue;
                    if (n >= CAND_CAP) return -1;
                    // insertion keeps the candidates sorted by index
                    int k = n++;
                    const int lj = j - base;
                    while (k > 0 && cand[k - 1] > lj) {
                        cand[k] = cand[k - 1];
                        k--;
                    }
                    cand[k] = lj;
                }
            }
    return n;
}

__global__ void __launch_bounds__(CELL_BLOCK)
    k_rebuild_cells(mte_params P, DevTopo T, DevState S, CellTables C, float radius, int K, int do_lj,
                    int do_bonded) {
    __shared__ int cand_all[CELL_BLOCK * CAND_CAP];
    const int b = blockIdx.x * CELL_BLOCK + threadIdx.x;
    if (b >= T.nb) return;
    const int t = T.tid[b], base = T.tbase[t];
    RebuildOwn o = rebuild_own(S.xyz[b], S.fn[b], S.fp1[b], S.fp2[b], T.mon_type[b], T.harm_raw[b], b - base, b,
                               base, T.tbox[t]);
    if (!T.extra[b]) {
        int *cand = cand_all + threadIdx.x * CAND_CAP;
        const int n = gather_candidates(T, S, C, b, radius, K, cand);
        if (n < 0) {
            o.ovf = 1;
        } else {
            for (int k = 0; k < n; k++) {
                const int j = base + cand[k];
                rebuild_pair(T, o, cand[k], S.xyz[j], S.fn[j], S.fp1[j], S.fp2[j], T.mon_type[j], P.ljpairscutoff,
                             do_lj, do_bonded);
            }
        }
    }
    rebuild_finish(T, o, do_lj, do_bonded);
}

float candidate_radius(const mte_params &P) {
    return rebuild_radius(P.ljpairscutoff, P.lj_on != 0, P.is_assembly != 0);   // superset of every list test
}

hipError_t build_cells(mte_engine *e, float side, int K) {
    const CellTables &C = e->cells;
    const int gb = (e->nb + 255) / 256;
    // the bucket counts are zero here: set at creation, and cleared by
    // k_scan_apply once it has read them
    hipLaunchKernelGGL(k_cell_assign, dim3(gb), dim3(256), 0, e->stream, e->T, e->S[e->cur], C, side, K);
    const int nblk = (C.total + CELL_SCAN_SPAN - 1) / CELL_SCAN_SPAN;
    hipLaunchKernelGGL(k_scan_partial, dim3(nblk), dim3(SCAN_T), 0, e->stream, C.count, C.total, C.bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, e->stream, C.bsum, nblk);
    hipLaunchKernelGGL(k_scan_apply, dim3(nblk), dim3(SCAN_T), 0, e->stream, C);
    hipLaunchKernelGGL(k_cell_scatter, dim3(gb), dim3(256), 0, e->stream, e->T, e->S[e->cur], C);
    return hipGetLastError();
}

hipError_t launch_rebuild_cells(mte_engine *e) {
    const int do_lj = e->P.lj_on ? 1 : 0;
    const int do_bonded = e->P.is_assembly ? 1 : 0;
    if (!do_lj && !do_bonded) return hipSuccess;
    const float radius = candidate_radius(e->P);
    const int K = e->tune.cell_stencil == 2 ? 2 : 1;
    hipError_t s = build_cells(e, radius / K, K);
    if (s != hipSuccess) return s;
    hipLaunchKernelGGL(k_rebuild_cells, dim3((e->nb + CELL_BLOCK - 1) / CELL_BLOCK), dim3(CELL_BLOCK), 0, e->stream,
                       e->P, e->T, e->S[e->cur], e->cells, radius, K, do_lj, do_bonded);
    return hipGetLastError();
}

}  // namespace mte
