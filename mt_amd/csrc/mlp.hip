// mlp.hip -- the learned hydrodynamic correction (MLP-HI); the reference has
// no counterpart.
//
// Per bead i the engine predicts the local hydrodynamic tensor
//   G_i = sum_{j != i, r_ij < 12 nm} phi(r_ij) D_hat_ij
// (D_hat the RPY block of src/bdhitea_kernel.cu:38-58, phi = 1 up to 8 nm and
// a cosine taper to 0 at 12 nm) from the relative positions of its 32 nearest
// neighbours (96 inputs), with a 96-128-128-128-9 tanh network, and replaces
// TEA's per-step O(N^2) pair sums by it (k_mlp_step), keeping TEA's beta and
// C_i normalisation.  The exact G_i (the same neighbour sum over every
// neighbour within 12 nm) is the training label and the sampled-error judge;
// plain TEA is the fallback.
//
// Kernels: neighbour features and exact labels over the hashed cell tables
// (cs >= 12 nm); inference fused with the integrator on the matrix cores
// (v_mfma_f32_16x16x32_bf16: bf16 weights, activations carried as bf16
// hi + lo pairs, FP32 accumulation; the weight fragments sit in LDS and the
// activations stay in registers from layer to layer); training (forward with
// stored activations, backward, Adam) in FP32 on the vector ALUs.
#include "engine_internal.h"

#include <mutex>

namespace mte {

constexpr int MLP_BLOCK = MLP_SEARCH_BLOCK;
constexpr int ROW = MLP_H + 1;   // LDS row pitch (odd: no bank conflicts between lanes)
constexpr float MLP_RC = 12.f, MLP_RIN = 8.f;
// flat parameter offsets
constexpr int OW1 = 0, OB1 = OW1 + MLP_IN * MLP_H, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H, OW3 = OB2 + MLP_H,
              OB3 = OW3 + MLP_H * MLP_H, OW4 = OB3 + MLP_H, OB4 = OW4 + MLP_H * MLP_OUT;
static_assert(OB4 + MLP_OUT == MLP_NPARAM, "parameter layout");

__device__ __forceinline__ float taper(float r) {
    if (r <= MLP_RIN) return 1.f;
    if (r >= MLP_RC) return 0.f;
    float s, c;
    sincos_fast(3.14159265358979f * (r - MLP_RIN) / (MLP_RC - MLP_RIN), &s, &c);
    return 0.5f * (1.f + c);
}

// ------------------------------------------------------------ features, labels
constexpr int FEAT_BLK_NN = 8;   // neighbours per block of written features (24 floats: whole
                                 // float4 stores and whole 8-float MFMA operand slices)
static_assert(3 * FEAT_BLK_NN % 8 == 0 && FEAT_BLK_NN % 4 == 0, "feature blocks");
// The 27 bucket ranges of a search as [first, end) slots of the bucket-ordered
// positions.  A trajectory's beads hold the slots [base, base + n) of its
// buckets, so with trajectory-local indices of 16 bits (n < 65536) both ends
// pack into one register.
template <typename JT>
struct BucketRanges {
    using W = int2;
    W v[27];
    __device__ __forceinline__ void set(int o, int2 r, int) { v[o] = make_int2(r.x, r.x + r.y); }
    __device__ __forceinline__ void clear(int o) { v[o] = make_int2(0, 0); }
    __device__ __forceinline__ static int lo(W w, int) { return w.x; }
    __device__ __forceinline__ static int hi(W w, int) { return w.y; }
};
template <>
struct BucketRanges<uint16_t> {
    using W = uint32_t;
    W v[27];
    __device__ __forceinline__ void set(int o, int2 r, int base) {
        const uint32_t s = (uint32_t)(r.x - base);
        v[o] = s | ((s + (uint32_t)r.y) << 16);
    }
    __device__ __forceinline__ void clear(int o) { v[o] = 0; }
    __device__ __forceinline__ static int lo(W w, int base) { return base + (int)(w & 0xffffu); }
    __device__ __forceinline__ static int hi(W w, int base) { return base + (int)(w >> 16); }
};

// Neighbour scan over the 27 cells of the hashed table (built with cell side
// >= 12 nm, or 3 cells per side covering the whole box).  The 27 bucket
// ranges are loaded up front; a bucket that two of the cells hash to is
// visited once, and the candidates are streamed from the bucket-ordered
// positions.  Every bead within 12 nm lies in one of the 27 cells, so
// visiting their buckets (each once) sees it exactly once, and the distance
// test rejects the beads of other cells sharing a bucket.  Keeps the 32
// nearest (distance, then index) and sums the exact label over every
// neighbour within 12 nm (LABELS: only on check steps and for training data).
// The 32 nearest live in registers as 64-bit keys (the squared distance's
// float bits, which order as integers for non-negative floats, over the
// trajectory-local index), ascending, empty slots all ones.  A candidate
// that precedes the last key is inserted without branches: its place is the
// count of smaller keys and every slot takes its own, its left neighbour's
// or the new key.  A wave's lanes then never wait on one another's shift
// loops, which on a crowded tube (60-130 candidates per lane, most of them
// inserted) had made a wave run the longest lane's shifts for every
// candidate.  JT selects the bucket-range packing (16-bit where every
// trajectory has fewer than 65536 beads).
// With fblk, a bead's features are written only up to the 24-float block
// (8 neighbours) holding its last neighbour, and the number of blocks goes
// to fblk: the step's network reads the rest as zeros, so in a dilute box
// (a few neighbours within 12 nm) most of the 384 feature bytes per bead
// never cross HBM.
template <bool LABELS, typename JT>
__global__ void __launch_bounds__(MLP_BLOCK)
    k_mlp_local(mte_params P, DevTopo T, DevState S, CellTables C, const int2 *blocks, float *feat, float *targ,
                uint8_t *fblk) {
    using Ranges = BucketRanges<JT>;
    // the lane's non-empty bucket ranges, in bucket order, for the walk over
    // its candidates
    __shared__ typename Ranges::W srg[27 * MLP_BLOCK];
    const int2 blk = blocks[blockIdx.x];
    const int t = blk.x;
    const int base = T.tbase[t];
    const int b = base + blk.y * MLP_BLOCK + threadIdx.x;
    if (b >= base + T.tsize[t]) return;
    uint64_t key[MLP_NN];
#pragma unroll
    for (int k = 0; k < MLP_NN; k++) key[k] = ~0ull;
    const float box = T.tbox[t], ibox = inv_side(box);
    const float3 ri = f3(S.xyz[b]);
    float A[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 0;
    if (!T.extra[b]) {
        const int dims = box > 0.f ? max(3, (int)floorf(box / MLP_RC)) : 0;
        const int4 me = C.cell[b];
        const uint32_t mask = (uint32_t)(C.tsize[t] - 1);
        const int toff = C.toff[t];
        const float ia = 1.f / P.tea_a;
        // own cell first, then faces, edges, corners: the nearest neighbours
        // come early and later candidates mostly fail the 32nd-distance test
        constexpr int ORD[27] = {13, 4, 10, 12, 14, 16, 22, 1, 3, 5, 7, 9, 11, 15, 17, 19, 21, 23, 25,
                                 0, 2, 6, 8, 18, 20, 24, 26};
        int bk[27];
#pragma unroll
        for (int o = 0; o < 27; o++) {
            const int c = ORD[o];
            int cx = me.x + c % 3 - 1, cy = me.y + (c / 3) % 3 - 1, cz = me.z + c / 9 - 1;
            if (dims > 0) {
                cx = wrap_cell(cx, dims);
                cy = wrap_cell(cy, dims);
                cz = wrap_cell(cz, dims);
            }
            const uint32_t h = ((uint32_t)cx * 73856093u) ^ ((uint32_t)cy * 19349663u) ^ ((uint32_t)cz * 83492791u);
            bk[o] = toff + (int)(h & mask);
        }
        Ranges rg;
#pragma unroll
        for (int o = 0; o < 27; o++) rg.set(o, C.range[bk[o]], base);
        // a bucket two cells hash to is visited once: a 64-bit filter on the
        // low bucket bits, and the full comparison only when it hits
        uint64_t seen = 1ull << (bk[0] & 63);
#pragma unroll
        for (int o = 1; o < 27; o++) {
            const uint64_t bit = 1ull << (bk[o] & 63);
            if (seen & bit) {
                bool dup = false;
#pragma unroll
                for (int k = 0; k < o; k++) dup |= bk[k] == bk[o];
                if (dup) rg.clear(o);
            }
            seen |= bit;
        }
        auto visit = [&](const float4 pj) {
            const int jw = __float_as_int(pj.w);
            const int j = jw & 0x7fffffff;
            const float3 d = min_image_pre(make_float3(pj.x - ri.x, pj.y - ri.y, pj.z - ri.z), box, ibox);
            const float r2 = d.x * d.x + d.y * d.y + d.z * d.z;
            // extra beads, this bead and beads beyond the radius drop out
            if (jw < 0 || j == b || r2 >= MLP_RC * MLP_RC) return;
            if (LABELS) {
                // exact label: phi(r) D_hat(d)
                const float r = sqrtf(r2), rinv = 1.f / r;
                const float ra = r * ia;
                float cii, crr;
                if (ra > 2.f) {
                    const float u = 1.f / ra, u2 = u * u;
                    cii = 0.75f * u * (1.f + (2.f / 3.f) * u2);
                    crr = 0.75f * u * (1.f - 2.f * u2);
                } else {
                    cii = 1.f - 9.f * ra / 32.f;
                    crr = 3.f * ra / 32.f;
                }
                const float w = taper(r);
                const float e[3] = {d.x * rinv, d.y * rinv, d.z * rinv};
#pragma unroll
                for (int x = 0; x < 3; x++)
#pragma unroll
                    for (int y = 0; y < 3; y++) A[3 * x + y] += w * (crr * e[x] * e[y] + (x == y ? cii : 0.f));
            }
            // 32 nearest, ordered by (r2, j)
            const uint64_t kn = ((uint64_t)__float_as_uint(r2) << 32) | (uint32_t)(j - base);
            if (kn < key[MLP_NN - 1]) {
                // slots in chunks of 8, a chunk only where some lane of the
                // wave holds a key in it (the rest are all ones and stay so):
                // a dilute wave touches one chunk.  Sorted keys: slot k keeps
                // its key if that is below kn, else takes kn when the key
                // below is below kn, else the key below; from the top down,
                // each comparison serves two slots.
                constexpr int CH = 8;
                bool used[MLP_NN / CH];
#pragma unroll
                for (int c = 0; c < MLP_NN / CH; c++) used[c] = c == 0 || __any(cnt >= CH * c);
#pragma unroll
                for (int c = MLP_NN / CH - 1; c >= 0; c--)
                    if (used[c]) {
#pragma unroll
                        for (int k = CH * (c + 1) - 1; k >= CH * c; k--) {
                            const bool keep = key[k] < kn;
                            const uint64_t below = k > 0 ? key[k - 1] : kn;
                            const bool below_lt = k > 0 && below < kn;
                            key[k] = keep ? key[k] : (below_lt ? kn : below);
                        }
                    }
                cnt = min(cnt + 1, MLP_NN);
            }
        };
        // the non-empty ranges, in bucket order, and their candidate count
        int nne = 0, total = 0;
#pragma unroll
        for (int o = 0; o < 27; o++) {
            const int q0 = Ranges::lo(rg.v[o], base), q1 = Ranges::hi(rg.v[o], base);
            if (q1 > q0) {
                srg[nne * MLP_BLOCK + threadIdx.x] = rg.v[o];
                nne++;
                total += q1 - q0;
            }
        }
        // One flat walk over the lane's candidates (buckets in order, each
        // bucket's in slot order), the loads PF candidates ahead of the
        // visits (pf[k % PF] holds candidate k; the cursor q, in range r,
        // is the next candidate to request).  A wave then runs as many
        // iterations as its lane with the most candidates, not the sum over
        // buckets of each bucket's longest lane.
        int r = 0, q = 0, qend = 0;
        if (nne > 0) {
            const typename Ranges::W w = srg[threadIdx.x];
            q = Ranges::lo(w, base);
            qend = Ranges::hi(w, base);
        }
        auto advance = [&]() {
            if (++q == qend && ++r < nne) {
                const typename Ranges::W w = srg[r * MLP_BLOCK + threadIdx.x];
                q = Ranges::lo(w, base);
                qend = Ranges::hi(w, base);
            }
        };
        constexpr int PF = 2;   // 3 and 4 measured no faster
        float4 pf[PF];
#pragma unroll
        for (int u = 0; u < PF; u++) pf[u] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int u = 0; u < PF; u++)
            if (u < total) {
                pf[u] = C.spos[q];
                advance();
            }
        for (int k = 0; k < total; k += PF) {
#pragma unroll
            for (int u = 0; u < PF; u++)
                if (k + u < total) {
                    const float4 pj = pf[u];
                    if (k + u + PF < total) {
                        pf[u] = C.spos[q];
                        advance();
                    }
                    visit(pj);
                }
        }
    }
    if (feat) {
        // a block of eight neighbours (24 floats, six float4 stores) at a
        // time, their positions requested together
        float4 *f = reinterpret_cast<float4 *>(feat + (size_t)b * MLP_IN);
        const int nblk = (cnt + FEAT_BLK_NN - 1) / FEAT_BLK_NN;
        if (fblk) fblk[b] = (uint8_t)nblk;
#pragma unroll
        for (int k0 = 0; k0 < MLP_NN; k0 += FEAT_BLK_NN) {
            if (fblk && k0 >= FEAT_BLK_NN * nblk) break;
            float4 xj[FEAT_BLK_NN];
#pragma unroll
            for (int u = 0; u < FEAT_BLK_NN; u++)
                xj[u] = k0 + u < cnt ? S.xyz[base + (int)(uint32_t)key[k0 + u]] : make_float4(0.f, 0.f, 0.f, 0.f);
            float v[3 * FEAT_BLK_NN];
#pragma unroll
            for (int u = 0; u < FEAT_BLK_NN; u++) {
                float3 d = make_float3(0.f, 0.f, 0.f);
                if (k0 + u < cnt) {
                    d = min_image_pre(sub3(f3(xj[u]), ri), box, ibox);
                    constexpr float irc = 1.f / MLP_RC;   // features in units of the radius
                    d = make_float3(d.x * irc, d.y * irc, d.z * irc);
                }
                v[3 * u] = d.x;
                v[3 * u + 1] = d.y;
                v[3 * u + 2] = d.z;
            }
#pragma unroll
            for (int w = 0; w < 3 * FEAT_BLK_NN / 4; w++)
                f[3 * k0 / 4 + w] = make_float4(v[4 * w], v[4 * w + 1], v[4 * w + 2], v[4 * w + 3]);
        }
    }
    if (LABELS) {
#pragma unroll
        for (int q = 0; q < 9; q++) targ[(size_t)b * MLP_OUT + q] = A[q];
    }
}

// ------------------------------------------------------------------ network
// One dense layer for this lane's sample: row (LDS, K inputs) -> row (M
// outputs), weights W[k * M + m] and bias, tanh when `act`.  Outputs are
// produced in chunks of 64 accumulators.
template <int K, int M>
__device__ __forceinline__ void dense_row(float *row, const float *__restrict__ W, const float *__restrict__ bias,
                                          bool act, float *out_regs) {
    constexpr int CH = M < 64 ? M : 64;
    float res[M];
#pragma unroll
    for (int m0 = 0; m0 < M; m0 += CH) {
        float acc[CH];
#pragma unroll
        for (int m = 0; m < CH; m++) acc[m] = bias[m0 + m];
        for (int k = 0; k < K; k++) {
            const float a = row[k];
            const float *w = W + k * M + m0;
#pragma unroll
            for (int m = 0; m < CH; m++) acc[m] = fmaf(a, w[m], acc[m]);
        }
#pragma unroll
        for (int m = 0; m < CH; m++) res[m0 + m] = act ? tanhf(acc[m]) : acc[m];
    }
    if (out_regs) {
#pragma unroll
        for (int m = 0; m < M; m++) out_regs[m] = res[m];
    } else {
#pragma unroll
        for (int m = 0; m < M; m++) row[m] = res[m];
    }
}

// Full network on the lane's LDS row (features in, 9 outputs to y).
// keep != nullptr stores the three hidden activations (training).
__device__ __forceinline__ void mlp_eval(float *row, const float *__restrict__ w, float y[MLP_OUT], float *keep,
                                         size_t stride) {
    dense_row<MLP_IN, MLP_H>(row, w + OW1, w + OB1, true, nullptr);
    if (keep)
        for (int m = 0; m < MLP_H; m++) keep[m] = row[m];
    dense_row<MLP_H, MLP_H>(row, w + OW2, w + OB2, true, nullptr);
    if (keep)
        for (int m = 0; m < MLP_H; m++) keep[stride + m] = row[m];
    dense_row<MLP_H, MLP_H>(row, w + OW3, w + OB3, true, nullptr);
    if (keep)
        for (int m = 0; m < MLP_H; m++) keep[2 * stride + m] = row[m];
    dense_row<MLP_H, MLP_OUT>(row, w + OW4, w + OB4, false, y);
}

// Cholesky factor of the symmetric 3x3 S (row-major); false when a pivot
// falls below 1e-3 (not positive definite enough to use).
__device__ __forceinline__ bool chol3(const float S[9], float Lc[6]) {
    const float l00s = S[0];
    if (!(l00s > 1e-3f)) return false;
    const float l00 = sqrtf(l00s);
    const float l10 = S[3] / l00, l20 = S[6] / l00;
    const float l11s = S[4] - l10 * l10;
    if (!(l11s > 1e-3f)) return false;
    const float l11 = sqrtf(l11s);
    const float l21 = (S[7] - l20 * l10) / l11;
    const float l22s = S[8] - l20 * l20 - l21 * l21;
    if (!(l22s > 1e-3f)) return false;
    Lc[0] = l00; Lc[1] = l10; Lc[2] = l11; Lc[3] = l20; Lc[4] = l21; Lc[5] = sqrtf(l22s);
    return true;
}

// ------------------------------------------------------ inference on MFMA
// Each layer is computed transposed, D^T = W^T X^T, with
// v_mfma_f32_16x16x32_bf16: output neurons are the 16 rows (M), 16 beads the
// columns (N), inputs the reduction (K = 32 per instruction).  Lane l supplies
// A[l%16][8(l/16)+q] and B[8(l/16)+q][l%16] (q = 0..7) and receives
// D[4(l/16)+r][l%16] (r = 0..3).  The outputs of two adjacent 16-neuron
// tiles therefore already sit in B-operand position for the next layer if
// that layer's K index 8g+q is read as neuron 4g+q (q < 4) or 16+4g+(q-4)
// (q >= 4) of the 32-neuron block: the weight image is packed in that order
// and activations never leave registers.  Activations enter as bf16 hi + lo
// pairs (|x - hi - lo| <= 2^-18 |x|), so with bf16 weights the products are
// those of the FP32 network with bf16-rounded weights, accumulated in FP32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int NET_WAVES = 12;                 // waves per workgroup (3 per SIMD)
constexpr int NET_THREADS = 64 * NET_WAVES;
constexpr int NET_GROUPS = 2;                 // 16-bead column groups per wave iteration
constexpr int NET_TILE = 16 * NET_GROUPS;     // beads per wave iteration
constexpr int MT_H = MLP_H / 16;              // output tiles of a hidden layer
constexpr int KS_IN = MLP_IN / 32, KS_H = MLP_H / 32;
// weight image in 16-byte fragments, [layer][m tile][k step][lane]
constexpr int FR1 = 0, FR2 = FR1 + MT_H * KS_IN * 64, FR3 = FR2 + MT_H * KS_H * 64, FR4 = FR3 + MT_H * KS_H * 64,
              FR_END = FR4 + KS_H * 64;
// FP32 biases after the fragments: b1, b2, b3, then b4 padded to 16
constexpr int BI2 = MLP_H, BI3 = 2 * MLP_H, BI4 = 3 * MLP_H, BI_END = 3 * MLP_H + 16;
constexpr size_t NET_IMG_BYTES = (size_t)FR_END * 16 + (size_t)BI_END * 4;
static_assert(MLP_IN % 32 == 0 && MLP_H % 32 == 0 && MLP_OUT <= 16, "network shape vs MFMA tiles");
static_assert(NET_IMG_BYTES % 16 == 0, "image size");

// Weight image from the bf16-rounded weights (exact bf16 values).
__global__ void __launch_bounds__(256) k_mlp_pack(const float *__restrict__ wi, u32x4 *img) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id < FR_END) {
        const float *W;
        int M, ks, mt, local;
        bool chained = true;
        if (id < FR2) {
            local = id - FR1; W = wi + OW1; M = MLP_H; chained = false;
            mt = (local / 64) / KS_IN; ks = (local / 64) % KS_IN;
        } else if (id < FR3) {
            local = id - FR2; W = wi + OW2; M = MLP_H;
            mt = (local / 64) / KS_H; ks = (local / 64) % KS_H;
        } else if (id < FR4) {
            local = id - FR3; W = wi + OW3; M = MLP_H;
            mt = (local / 64) / KS_H; ks = (local / 64) % KS_H;
        } else {
            local = id - FR4; W = wi + OW4; M = MLP_OUT;
            mt = 0; ks = local / 64;
        }
        const int lane = local % 64, g = lane / 16, m = 16 * mt + lane % 16;
        unsigned int word[4] = {0u, 0u, 0u, 0u};
        for (int q = 0; q < 8; q++) {
            const int k = 32 * ks + (chained ? (q < 4 ? 4 * g + q : 16 + 4 * g + (q - 4)) : 8 * g + q);
            const float v = m < M ? W[k * M + m] : 0.f;
            word[q / 2] |= (__float_as_uint(v) >> 16) << (16 * (q % 2));
        }
        img[id] = u32x4{word[0], word[1], word[2], word[3]};
    } else if (id < FR_END + BI_END) {
        const int j = id - FR_END;
        float v;
        if (j < BI2) v = wi[OB1 + j];
        else if (j < BI3) v = wi[OB2 + j - BI2];
        else if (j < BI4) v = wi[OB3 + j - BI3];
        else v = j - BI4 < MLP_OUT ? wi[OB4 + j - BI4] : 0.f;
        reinterpret_cast<float *>(img + FR_END)[j] = v;
    }
}

__device__ __forceinline__ float tanh_fast(float x) {
    // tanh x = 1 - 2 / (exp(2x) + 1) for either sign: absolute error a few
    // 1e-7; exp(2x) overflows to inf for large x (1) and underflows to 0
    // for large -x (-1)
    const float e = __expf(2.f * x);
    return fmaf(-2.f, rcp_fast(e + 1.f), 1.f);
}

__device__ __forceinline__ void split8(const float v[8], bf16x8 &hi, bf16x8 &lo) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const __bf16 h = (__bf16)v[q];
        hi[q] = h;
        lo[q] = (__bf16)(v[q] - (float)h);
    }
}

// One layer over the bead groups: acc[mt][grp] = bias + W^T (hi + lo).
template <int KS, int MT>
__device__ __forceinline__ void net_layer(const u32x4 *frag, const float *bias, const bf16x8 (&hi)[KS][NET_GROUPS],
                                          const bf16x8 (&lo)[KS][NET_GROUPS], f32x4 (&acc)[MT][NET_GROUPS], int lane) {
    const int g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < MT; mt++) {
        const f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 16 * mt + 4 * g);
#pragma unroll
        for (int grp = 0; grp < NET_GROUPS; grp++) acc[mt][grp] = b;
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, frag[(mt * KS + ks) * 64 + lane]);
#pragma unroll
            for (int grp = 0; grp < NET_GROUPS; grp++) {
                acc[mt][grp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hi[ks][grp], acc[mt][grp], 0, 0, 0);
                acc[mt][grp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lo[ks][grp], acc[mt][grp], 0, 0, 0);
            }
        }
    }
}

// The input layer over the k-slices in kuse (bit ks: features 32 ks ..
// 32 ks + 31); a slice outside kuse is zero for every bead of the tile and
// adds nothing.
__device__ __forceinline__ void net_layer_in(const u32x4 *frag, const float *bias,
                                             const bf16x8 (&hi)[KS_IN][NET_GROUPS],
                                             const bf16x8 (&lo)[KS_IN][NET_GROUPS], f32x4 (&acc)[MT_H][NET_GROUPS],
                                             int lane, uint32_t kuse) {
    const int g = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < MT_H; mt++) {
        const f32x4 b = *reinterpret_cast<const f32x4 *>(bias + 16 * mt + 4 * g);
#pragma unroll
        for (int grp = 0; grp < NET_GROUPS; grp++) acc[mt][grp] = b;
    }
#pragma unroll
    for (int ks = 0; ks < KS_IN; ks++) {
        if (!((kuse >> ks) & 1u)) continue;
#pragma unroll
        for (int mt = 0; mt < MT_H; mt++) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, frag[(mt * KS_IN + ks) * 64 + lane]);
#pragma unroll
            for (int grp = 0; grp < NET_GROUPS; grp++) {
                acc[mt][grp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, hi[ks][grp], acc[mt][grp], 0, 0, 0);
                acc[mt][grp] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, lo[ks][grp], acc[mt][grp], 0, 0, 0);
            }
        }
    }
}

// tanh of a hidden layer, repacked as the next layer's B fragments.
template <int KS>
__device__ __forceinline__ void hidden_to_frags(f32x4 (&acc)[2 * KS][NET_GROUPS], bf16x8 (&hi)[KS][NET_GROUPS],
                                                bf16x8 (&lo)[KS][NET_GROUPS]) {
#pragma unroll
    for (int ks = 0; ks < KS; ks++)
#pragma unroll
        for (int grp = 0; grp < NET_GROUPS; grp++) {
            float v[8];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                v[r] = tanh_fast(acc[2 * ks][grp][r]);
                v[4 + r] = tanh_fast(acc[2 * ks + 1][grp][r]);
            }
            split8(v, hi[ks][grp], lo[ks][grp]);
        }
}

// The network on beads b0 .. b0+NET_TILE-1 (features of beads >= n read as
// zero); returns in lane t < NET_TILE the 9 outputs of bead b0 + t.
// lim, byte grp: the features of the lane's bead in column group grp past
// that count are zeros and are not read (feat_limits).
__device__ __forceinline__ void net_tile(const u32x4 *frag, const float *bias, const float *__restrict__ feat,
                                         uint32_t lim, int b0, int lane, float y[MLP_OUT]) {
    const int g = lane >> 4, col = lane & 15;
    // The fragments are the same for every tile: keep the compiler from
    // hoisting all of them (368 registers) out of the caller's tile loop.
    __asm__ volatile("" ::: "memory");
    // k-slices of the input with a feature present for some bead of the tile
    int most = 0;
#pragma unroll
    for (int grp = 0; grp < NET_GROUPS; grp++) most = max(most, (int)((lim >> (8 * grp)) & 0xffu));
    uint32_t kuse = 0;
#pragma unroll
    for (int ks = 0; ks < KS_IN; ks++) kuse |= __any(most > 32 * ks) ? 1u << ks : 0u;
    bf16x8 hi1[KS_IN][NET_GROUPS], lo1[KS_IN][NET_GROUPS];
#pragma unroll
    for (int grp = 0; grp < NET_GROUPS; grp++) {
        const int bead = b0 + 16 * grp + col;
#pragma unroll
        for (int ks = 0; ks < KS_IN; ks++) {
            if (!((kuse >> ks) & 1u)) continue;
            float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (32 * ks + 8 * g < (int)((lim >> (8 * grp)) & 0xffu)) {
                const float4 *src = reinterpret_cast<const float4 *>(feat + (size_t)bead * MLP_IN + 32 * ks + 8 * g);
                const float4 u = src[0], w = src[1];
                v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
                v[4] = w.x; v[5] = w.y; v[6] = w.z; v[7] = w.w;
            }
            split8(v, hi1[ks][grp], lo1[ks][grp]);
        }
    }
    f32x4 acc[MT_H][NET_GROUPS];
    net_layer_in(frag + FR1, bias, hi1, lo1, acc, lane, kuse);
    bf16x8 hi[KS_H][NET_GROUPS], lo[KS_H][NET_GROUPS];
    hidden_to_frags<KS_H>(acc, hi, lo);
    net_layer<KS_H, MT_H>(frag + FR2, bias + BI2, hi, lo, acc, lane);
    hidden_to_frags<KS_H>(acc, hi, lo);
    net_layer<KS_H, MT_H>(frag + FR3, bias + BI3, hi, lo, acc, lane);
    hidden_to_frags<KS_H>(acc, hi, lo);
    f32x4 out[1][NET_GROUPS];
    net_layer<KS_H, 1>(frag + FR4, bias + BI4, hi, lo, out, lane);
    // output q of bead 16 grp + c sits in lane 16 (q / 4) + c, element q % 4
    const int grp = (lane >> 4) % NET_GROUPS;
#pragma unroll
    for (int q = 0; q < MLP_OUT; q++) {
        const int src = 16 * (q >> 2) + col;
        float v = 0.f;
#pragma unroll
        for (int gg = 0; gg < NET_GROUPS; gg++) {
            const float vg = __shfl(out[0][gg][q & 3], src);
            v = gg == grp ? vg : v;
        }
        y[q] = v;
    }
}

// Features present for the lane's beads of the tile at b0, one byte per
// column group (net_tile's lim): all 96 without fblk, else fblk's blocks of
// 24; none past the last bead.
__device__ __forceinline__ uint32_t feat_limits(const uint8_t *__restrict__ fblk, int n, int b0, int lane) {
    uint32_t lim = 0;
#pragma unroll
    for (int grp = 0; grp < NET_GROUPS; grp++) {
        const int bead = b0 + 16 * grp + (lane & 15);
        const int l = bead >= n ? 0 : fblk ? 3 * FEAT_BLK_NN * (int)fblk[bead] : MLP_IN;
        lim |= (uint32_t)l << (8 * grp);
    }
    return lim;
}
static_assert(MLP_IN < 256 && NET_GROUPS <= 4, "feature limits packed in bytes");

// Weight image into LDS (dynamic, NET_IMG_BYTES).
__device__ __forceinline__ void load_image(const u32x4 *__restrict__ img, u32x4 *lds) {
    for (int i = threadIdx.x; i < (int)(NET_IMG_BYTES / 16); i += NET_THREADS) lds[i] = img[i];
    __syncthreads();
}

__global__ void __launch_bounds__(NET_THREADS)
    k_mlp_forward(const u32x4 *__restrict__ img, const float *__restrict__ feat, int n, float *out) {
    extern __shared__ u32x4 net_lds[];
    load_image(img, net_lds);
    const float *bias = reinterpret_cast<const float *>(net_lds + FR_END);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntiles = (n + NET_TILE - 1) / NET_TILE;
    for (int tile = blockIdx.x * NET_WAVES + wave; tile < ntiles; tile += gridDim.x * NET_WAVES) {
        float y[MLP_OUT];
        net_tile(net_lds, bias, feat, feat_limits(nullptr, n, tile * NET_TILE, lane), tile * NET_TILE, lane, y);
        const int s = tile * NET_TILE + lane;
        if (lane < NET_TILE && s < n)
            for (int q = 0; q < MLP_OUT; q++) out[(size_t)s * MLP_OUT + q] = y[q];
    }
}

struct MlpStepArgs {
    mte_params P;
    DevTopo T;
    DevState Sin, Sout;
    const float4 *fxyz, *fang, *ci;
    const float *beta;
    uint4 *seeds;
    const float *noise;
    float *yout;
    int *nbad;
    LatOffsets L;
};

// MLP-HI integration of bead b from the network outputs y.  With G the
// predicted local HI tensor (symmetrised), TEA's per-trajectory beta and
// per-bead row normalisation C_i (refreshed by the epsilon pass as in TEA):
//   drift  (dt/gammaR) (I + G) F_i               local form of F_i + sum_j D_hat_ij F_j
//   noise  varR L xi,  L L^T = C (I + 2 beta G) C  row sum of TEA's noise covariance
//                                                  (src/bdhitea_kernel.cu:104-213) to first order in beta
// so cluster centre-of-mass diffusion follows TEA's.  A prediction whose
// covariance is not positive definite falls back to C^2 for the bead (counted).
// Noise draws as the plain integrator (mobile beads only).
// What the integrator reads of bead b, requested before the tile's network
// runs so that the loads complete behind the matrix work.
struct MlpIn {
    float4 x, an, n4, p1, p2, F, fa, cc;
    uint4 s0, s1;
    float bt;
    bool mobile;
};
__device__ __forceinline__ MlpIn mlp_load(const MlpStepArgs &A, int b) {
    MlpIn v;
    const DevTopo &T = A.T;
    v.x = A.Sin.xyz[b];
    v.an = A.Sin.ang[b];
    v.n4 = A.Sin.fn[b];
    v.p1 = A.Sin.fp1[b];
    v.p2 = A.Sin.fp2[b];
    v.F = A.fxyz[b];
    v.fa = A.fang[b];
    v.cc = A.ci[b];
    v.bt = A.beta[T.tid[b]];
    v.mobile = !T.fixed[b] && !T.extra[b];
    if (!A.noise) {
        v.s0 = A.seeds[b];
        v.s1 = A.seeds[b + T.nb];
    }
    return v;
}

__device__ __forceinline__ void mlp_integrate(const MlpStepArgs &A, int b, const float y[MLP_OUT], const MlpIn &in) {
    const mte_params &P = A.P;
    const DevTopo &T = A.T;
    float4 x = in.x, an = in.an;
    float4 n4 = in.n4, p1 = in.p1, p2 = in.p2;
    const bool mobile = in.mobile;
    if (mobile) {
        if (A.yout)
            for (int q = 0; q < MLP_OUT; q++) A.yout[(size_t)b * MLP_OUT + q] = y[q];
        float G[9];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) G[3 * r + c] = 0.5f * (y[3 * r + c] + y[3 * c + r]);
        const float bt = in.bt;
        const float4 cc = in.cc;
        const float Cd[3] = {rsq_fast(1.f + bt * bt * cc.x), rsq_fast(1.f + bt * bt * cc.y),
                             rsq_fast(1.f + bt * bt * cc.z)};
        float Sg[9], Lc[6];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) Sg[3 * r + c] = Cd[r] * Cd[c] * ((r == c ? 1.f : 0.f) + 2.f * bt * G[3 * r + c]);
        if (!chol3(Sg, Lc)) {
            atomicAdd(A.nbad, 1);
            Lc[0] = Cd[0]; Lc[1] = 0.f; Lc[2] = Cd[1]; Lc[3] = 0.f; Lc[4] = 0.f; Lc[5] = Cd[2];
        }
        float4 rx, ra;
        if (A.noise) {
            const float *nz = A.noise + 6 * (size_t)b;
            rx = make_float4(nz[0], nz[1], nz[2], 0.f);
            ra = make_float4(nz[3], nz[4], nz[5], 0.f);
        } else {
            uint4 s = in.s0;
            rx = normals4(s);
            A.seeds[b] = s;
            s = in.s1;
            ra = normals4(s);
            A.seeds[b + T.nb] = s;
        }
        const float4 F = in.F;
        const float dtg = P.dt / P.gammaR;
        const float d0 = F.x + G[0] * F.x + G[1] * F.y + G[2] * F.z;
        const float d1 = F.y + G[3] * F.x + G[4] * F.y + G[5] * F.z;
        const float d2 = F.z + G[6] * F.x + G[7] * F.y + G[8] * F.z;
        const float n0 = Lc[0] * rx.x, n1 = Lc[1] * rx.x + Lc[2] * rx.y, n2 = Lc[3] * rx.x + Lc[4] * rx.y + Lc[5] * rx.z;
        x.x += dtg * d0 + P.varR * n0;
        x.y += dtg * d1 + P.varR * n1;
        x.z += dtg * d2 + P.varR * n2;
        const float4 fa = in.fa;
        const float dta = P.dt / (P.gammaTheta * P.alpha), dtt = P.dt / P.gammaTheta;
        const float vta = P.varTheta * sqrtf(P.freeze_temp / P.alpha);
        an.x += dta * fa.x + vta * ra.x;
        an.z += dta * fa.z + vta * ra.y;
        an.y += dtt * fa.y + P.varTheta * ra.z;
        frame_vectors(an.x, an.y, an.z, A.L, n4, p1, p2);
    } else if (A.yout) {
        for (int q = 0; q < MLP_OUT; q++) A.yout[(size_t)b * MLP_OUT + q] = 0.f;
    }
    A.Sout.xyz[b] = x;
    A.Sout.ang[b] = an;
    A.Sout.fn[b] = n4;
    A.Sout.fp1[b] = p1;
    A.Sout.fp2[b] = p2;
}

// MLP-HI step: network on the matrix cores, then lanes 0..NET_TILE-1
// integrate the tile's beads.
__global__ void __launch_bounds__(NET_THREADS)
    k_mlp_step(const u32x4 *__restrict__ img, const float *__restrict__ feat, const uint8_t *__restrict__ fblk,
               MlpStepArgs A) {
    extern __shared__ u32x4 net_lds[];
    load_image(img, net_lds);
    const float *bias = reinterpret_cast<const float *>(net_lds + FR_END);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = A.T.nb;
    const int ntiles = (n + NET_TILE - 1) / NET_TILE;
    for (int tile = blockIdx.x * NET_WAVES + wave; tile < ntiles; tile += gridDim.x * NET_WAVES) {
        const uint32_t lim = feat_limits(fblk, n, tile * NET_TILE, lane);
        const int b = tile * NET_TILE + lane;
        const bool integ = lane < NET_TILE && b < n;
        float y[MLP_OUT];
        net_tile(net_lds, bias, feat, lim, tile * NET_TILE, lane, y);
        if (integ) mlp_integrate(A, b, y, mlp_load(A, b));
    }
}

// Per-trajectory squared error and squared label norm of predictions y
// against labels t over mobile beads (sampled-error judge), FP64 sums: each
// of MLP_ERR_PARTS workgroups per trajectory sums every MLP_ERR_PARTS-th
// 256-bead slice into its partial, and k_mlp_error_sum adds the partials in
// order.
__global__ void __launch_bounds__(256)
    k_mlp_error(DevTopo T, const float *y, const float *t, double *part) {
    __shared__ double se[256], sn[256];
    const int tr = blockIdx.x, p = blockIdx.y;
    const int base = T.tbase[tr], n = T.tsize[tr];
    double e = 0, m = 0;
    for (int i = p * 256 + threadIdx.x; i < n; i += 256 * MLP_ERR_PARTS) {
        const int b = base + i;
        if (T.fixed[b] || T.extra[b]) continue;
        for (int q = 0; q < MLP_OUT; q++) {
            const double d = (double)y[(size_t)b * MLP_OUT + q] - (double)t[(size_t)b * MLP_OUT + q];
            e += d * d;
            m += (double)t[(size_t)b * MLP_OUT + q] * (double)t[(size_t)b * MLP_OUT + q];
        }
    }
    se[threadIdx.x] = e;
    sn[threadIdx.x] = m;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            se[threadIdx.x] += se[threadIdx.x + k];
            sn[threadIdx.x] += sn[threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * ((size_t)tr * MLP_ERR_PARTS + p)] = se[0];
        part[2 * ((size_t)tr * MLP_ERR_PARTS + p) + 1] = sn[0];
    }
}

__global__ void __launch_bounds__(64) k_mlp_error_sum(int ntraj, const double *part, double *out) {
    const int tr = blockIdx.x * 64 + threadIdx.x;
    if (tr >= ntraj) return;
    double e = 0, m = 0;
    for (int p = 0; p < MLP_ERR_PARTS; p++) {
        e += part[2 * ((size_t)tr * MLP_ERR_PARTS + p)];
        m += part[2 * ((size_t)tr * MLP_ERR_PARTS + p) + 1];
    }
    out[2 * tr] = e;
    out[2 * tr + 1] = m;
}

// ----------------------------------------------------------------- training
__global__ void __launch_bounds__(MLP_BLOCK)
    k_mlp_train_fwd(const float *__restrict__ w, const float *feat, int n, float *hid, float *y) {
    __shared__ float rows[MLP_BLOCK * ROW];
    const int s = blockIdx.x * MLP_BLOCK + threadIdx.x;
    if (s >= n) return;
    float *row = rows + threadIdx.x * ROW;
    for (int k = 0; k < MLP_IN; k++) row[k] = feat[(size_t)s * MLP_IN + k];
    float out[MLP_OUT];
    // hidden layers stored sample-major: hid[(layer * n + s) * H + m]
    mlp_eval(row, w, out, hid + (size_t)s * MLP_H, (size_t)n * MLP_H);
    for (int q = 0; q < MLP_OUT; q++) y[(size_t)s * MLP_OUT + q] = out[q];
}

// dL/dy = 2 (y - t) / (n * 9); per-block squared-error partials.
__global__ void __launch_bounds__(256)
    k_mlp_out_grad(const float *y, const float *t, int n, float *g, double *partial) {
    __shared__ double red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double e = 0;
    if (i < n * MLP_OUT) {
        const float d = y[i] - t[i];
        g[i] = 2.f * d / (float)(n * MLP_OUT);
        e = (double)d * d;
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// dW[k][m] = sum_s in[s][k] g[s][m]; row k == K holds db[m] = sum_s g[s][m].
// blockIdx.y splits the samples into DW_SPLIT slices added atomically into
// zeroed dW / db.
constexpr int DW_SPLIT = 32;
__global__ void __launch_bounds__(256)
    k_mlp_dw(const float *in, const float *g, int n, int K, int M, float *dW, float *db) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= (K + 1) * M) return;
    const int k = id / M, m = id % M;
    const int per = (n + DW_SPLIT - 1) / DW_SPLIT;
    const int s0 = blockIdx.y * per, s1 = min(n, s0 + per);
    float acc = 0.f;
    if (k < K) {
        for (int s = s0; s < s1; s++) acc = fmaf(in[(size_t)s * K + k], g[(size_t)s * M + m], acc);
        atomicAdd(&dW[k * M + m], acc);
    } else {
        for (int s = s0; s < s1; s++) acc += g[(size_t)s * M + m];
        atomicAdd(&db[m], acc);
    }
}

// gin[s][k] = (sum_m g[s][m] W[k][m]) (1 - h[s][k]^2)   (tanh derivative)
__global__ void __launch_bounds__(256)
    k_mlp_dx(const float *g, const float *__restrict__ W, const float *h, int n, int K, int M, float *gin) {
    const int id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n * K) return;
    const int s = id / K, k = id % K;
    float acc = 0.f;
    for (int m = 0; m < M; m++) acc = fmaf(g[(size_t)s * M + m], W[k * M + m], acc);
    const float hv = h[(size_t)s * K + k];
    gin[id] = acc * (1.f - hv * hv);
}

__device__ __forceinline__ float round_bf16(float v) {
    uint32_t u = __float_as_uint(v);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(u);
}

// Adam (Kingma & Ba 2015) on the FP32 master weights; refreshes the
// bf16-rounded copy used by inference.
__global__ void __launch_bounds__(256)
    k_mlp_adam(float *w, const float *g, float *m1, float *m2, float *wi, int n, float lr, float b1, float b2,
               float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    const float nw = w[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
    w[i] = nw;
    wi[i] = round_bf16(nw);
}

__global__ void __launch_bounds__(256) k_fill_f32(float *p, int n, float v) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_mlp_round(const float *w, float *wi, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) wi[i] = round_bf16(w[i]);
}

// ------------------------------------------------------------------ launchers
static inline int blocks(int n, int bs) { return (n + bs - 1) / bs; }

hipError_t launch_mlp_cells(mte_engine *e) { return build_cells(e, MLP_RC, 1); }

hipError_t launch_mlp_local(mte_engine *e, float *feat, float *targ, bool build_tables, uint8_t *fblk) {
    if (build_tables) {
        const hipError_t s = launch_mlp_cells(e);
        if (s != hipSuccess) return s;
    }
    const dim3 grid(e->mlp.nblocks), block(MLP_BLOCK);
    const int2 *bl = e->mlp.blocks;
    const bool narrow = e->nmax < 65536;   // trajectory-local indices and bucket ends fit 16 bits
    if (targ && narrow)
        hipLaunchKernelGGL((k_mlp_local<true, uint16_t>), grid, block, 0, e->stream, e->P, e->T, e->S[e->cur],
                           e->cells, bl, feat, targ, fblk);
    else if (targ)
        hipLaunchKernelGGL((k_mlp_local<true, int>), grid, block, 0, e->stream, e->P, e->T, e->S[e->cur], e->cells,
                           bl, feat, targ, fblk);
    else if (narrow)
        hipLaunchKernelGGL((k_mlp_local<false, uint16_t>), grid, block, 0, e->stream, e->P, e->T, e->S[e->cur],
                           e->cells, bl, feat, targ, fblk);
    else
        hipLaunchKernelGGL((k_mlp_local<false, int>), grid, block, 0, e->stream, e->P, e->T, e->S[e->cur], e->cells,
                           bl, feat, targ, fblk);
    return hipGetLastError();
}

// Workgroups of the inference kernels: one per CU at most (the weight image
// takes most of a CU's LDS), each wave looping over bead tiles.
static int net_blocks(const mte_engine *e, int n) {
    const int need = blocks(blocks(n, NET_TILE), NET_WAVES);
    return std::max(1, std::min(need, e->num_cus));
}

static hipError_t net_lds_attr(const void *k) {
    return hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)NET_IMG_BYTES);
}

hipError_t launch_mlp_forward(mte_engine *e, const float *feat, int n, float *out) {
    hipError_t s = net_lds_attr((const void *)k_mlp_forward);
    if (s != hipSuccess) return s;
    hipLaunchKernelGGL(k_mlp_forward, dim3(net_blocks(e, n)), dim3(NET_THREADS), NET_IMG_BYTES, e->stream,
                       reinterpret_cast<const u32x4 *>(e->mlp.img), feat, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_step(mte_engine *e, const float *noise_step, float *yout) {
    hipError_t s = net_lds_attr((const void *)k_mlp_step);
    if (s != hipSuccess) return s;
    MlpStepArgs A{e->P, e->T, e->S[e->cur], e->S[1 - e->cur], e->fxyz, e->fang, e->tea.ci, e->tea.beta,
                  e->seeds, noise_step, yout, e->mlp.nbad, e->L};
    hipLaunchKernelGGL(k_mlp_step, dim3(net_blocks(e, e->nb)), dim3(NET_THREADS), NET_IMG_BYTES, e->stream,
                       reinterpret_cast<const u32x4 *>(e->mlp.img), e->mlp.feat, e->mlp.fblk, A);
    e->cur = 1 - e->cur;
    return hipGetLastError();
}

hipError_t launch_mlp_error(mte_engine *e, const float *y, const float *t) {
    hipLaunchKernelGGL(k_mlp_error, dim3(e->ntraj, MLP_ERR_PARTS), dim3(256), 0, e->stream, e->T, y, t,
                       e->mlp.err_part);
    hipLaunchKernelGGL(k_mlp_error_sum, dim3((e->ntraj + 63) / 64), dim3(64), 0, e->stream, e->ntraj,
                       (const double *)e->mlp.err_part, e->mlp.err);
    return hipGetLastError();
}

// Forward + backward over n samples already on the device; gradient into
// e->mlp.grad, squared-error partials into e->mlp.partial.
hipError_t launch_mlp_grad(mte_engine *e, const float *feat, const float *targ, int n) {
    MlpBuffers &M = e->mlp;
    const size_t H = MLP_H;
    float *h1 = M.hid, *h2 = M.hid + (size_t)n * H, *h3 = M.hid + 2 * (size_t)n * H;
    hipLaunchKernelGGL(k_mlp_train_fwd, dim3(blocks(n, MLP_BLOCK)), dim3(MLP_BLOCK), 0, e->stream, M.w, feat, n,
                       M.hid, M.y);
    hipLaunchKernelGGL(k_mlp_out_grad, dim3(blocks(n * MLP_OUT, 256)), dim3(256), 0, e->stream, M.y, targ, n, M.g4,
                       M.partial);
    float *gw = M.grad;
    hipLaunchKernelGGL(k_fill_f32, dim3(blocks(MLP_NPARAM, 256)), dim3(256), 0, e->stream, gw, MLP_NPARAM, 0.f);
    hipLaunchKernelGGL(k_mlp_dw, dim3(blocks((MLP_H + 1) * MLP_OUT, 256), DW_SPLIT), dim3(256), 0, e->stream, h3,
                       M.g4, n, MLP_H, MLP_OUT, gw + OW4, gw + OB4);
    hipLaunchKernelGGL(k_mlp_dx, dim3(blocks(n * MLP_H, 256)), dim3(256), 0, e->stream, M.g4, M.w + OW4, h3, n,
                       MLP_H, MLP_OUT, M.ga);
    hipLaunchKernelGGL(k_mlp_dw, dim3(blocks((MLP_H + 1) * MLP_H, 256), DW_SPLIT), dim3(256), 0, e->stream, h2, M.ga,
                       n, MLP_H, MLP_H, gw + OW3, gw + OB3);
    hipLaunchKernelGGL(k_mlp_dx, dim3(blocks(n * MLP_H, 256)), dim3(256), 0, e->stream, M.ga, M.w + OW3, h2, n,
                       MLP_H, MLP_H, M.gb);
    hipLaunchKernelGGL(k_mlp_dw, dim3(blocks((MLP_H + 1) * MLP_H, 256), DW_SPLIT), dim3(256), 0, e->stream, h1, M.gb,
                       n, MLP_H, MLP_H, gw + OW2, gw + OB2);
    hipLaunchKernelGGL(k_mlp_dx, dim3(blocks(n * MLP_H, 256)), dim3(256), 0, e->stream, M.gb, M.w + OW2, h1, n,
                       MLP_H, MLP_H, M.ga);
    hipLaunchKernelGGL(k_mlp_dw, dim3(blocks((MLP_IN + 1) * MLP_H, 256), DW_SPLIT), dim3(256), 0, e->stream, feat,
                       M.ga, n, MLP_IN, MLP_H, gw + OW1, gw + OB1);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(mte_engine *e, float lr, int t) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_mlp_adam, dim3(blocks(MLP_NPARAM, 256)), dim3(256), 0, e->stream, e->mlp.w, e->mlp.grad,
                       e->mlp.m1, e->mlp.m2, e->mlp.wi, MLP_NPARAM, lr, b1, b2, 1e-8f, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_round(mte_engine *e) {
    hipLaunchKernelGGL(k_mlp_round, dim3(blocks(MLP_NPARAM, 256)), dim3(256), 0, e->stream, e->mlp.w, e->mlp.wi,
                       MLP_NPARAM);
    return launch_mlp_pack(e);
}

hipError_t launch_mlp_pack(mte_engine *e) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(blocks(FR_END + BI_END, 256)), dim3(256), 0, e->stream, e->mlp.wi,
                       reinterpret_cast<u32x4 *>(e->mlp.img));
    return hipGetLastError();
}

size_t mlp_image_bytes() { return NET_IMG_BYTES; }

// ------------------------------------------- inference without allocation
// Staging of mte_mlp_forward_noalloc: features (MLP_NOALLOC_ROWS x 96) then
// outputs (MLP_NOALLOC_ROWS x 9), in one fixed device array.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code
// never uses the array's name as an address: the per-device address comes
// from hipGetSymbolAddress below.
static __device__ __align__(256) float scratch[(size_t)MLP_NOALLOC_ROWS * (MLP_IN + MLP_OUT)];

namespace {
std::mutex g_scratch_mu;
float *g_scratch_ptr[MLP_NOALLOC_MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device, float **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= MLP_NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {   // once per device, with that device current
        float *p = nullptr;
        const hipError_t s = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
        if (s != hipSuccess) return s;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace mte
