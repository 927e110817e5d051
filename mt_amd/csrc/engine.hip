// engine.hip -- the C ABI of include/mt_engine.h: device state, topology
// upload, the per-step driver and host<->device exchange.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "engine_internal.h"

using namespace mte;

static thread_local std::string g_create_error;

#define HIPCHK(call)                                                                           \
    do {                                                                                       \
        hipError_t _s = (call);                                                                \
        if (_s != hipSuccess) {                                                                \
            e->err = std::string(#call) + ": " + hipGetErrorString(_s);                       \
            return MTE_ERR_HIP;                                                                \
        }                                                                                      \
    } while (0)

template <class T>
static hipError_t dalloc(mte_engine *e, T **p, size_t count) {
    void *q = nullptr;
    hipError_t s = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (s == hipSuccess) {
        e->allocations.push_back(q);
        *p = static_cast<T *>(q);
    }
    return s;
}

static int fail(mte_engine *e, int code, const std::string &msg) {
    e->err = msg;
    return code;
}

static void release(mte_engine *e, void *p) {
    if (!p) return;
    (void)hipFree(p);
    e->allocations.erase(std::remove(e->allocations.begin(), e->allocations.end(), p), e->allocations.end());
}

// Work units (trajectory, 256-bead i group, j chunk) of the TEA pair kernel
// and its per-chunk partial buffer, for j-chunks of `chunk` beads; for
// MTE_TEA_SYMMETRIC and MTE_TEA_MFMA, units (trajectory, 64-bead chunk I,
// chunk J >= I) (for MTE_TEA_SYMMETRIC, one or two I chunks below J), and
// for MTE_TEA_MFMA the tile tables as well.
hipError_t mte::tea_build_units(mte_engine *e, int chunk) {
    release(e, e->tea.units);
    release(e, e->tea.nchunks);
    release(e, e->tea.partial);
    release(e, e->tea.tiles);
    release(e, e->tea.mom);
    release(e, e->tea.ctr);
    release(e, e->tea.rec);
    release(e, e->tea.rbead);
    e->tea.units = nullptr;
    e->tea.nchunks = nullptr;
    e->tea.partial = nullptr;
    e->tea.tiles = nullptr;
    e->tea.mom = nullptr;
    e->tea.ctr = nullptr;
    e->tea.rec = nullptr;
    e->tea.rbead = nullptr;
    e->tea.ntiles = 0;
    std::vector<int4> units;
    std::vector<int> nchunks(e->ntraj, 0);
    int maxc = 0;
    for (int t = 0; t < e->ntraj; t++) {
        const int n = (int)(e->off[t + 1] - e->off[t]);
        int c = 0;
        if (chunk == MTE_TEA_SYMMETRIC) {
            // (trajectory, I, J, number of I chunks): the diagonal chunk
            // alone, the chunks below J two at a time (I and I + 64)
            for (int j0 = 0; j0 < n; j0 += 64, c++) {
                units.push_back(make_int4(t, j0, j0, 1));
                for (int i0 = 0; i0 < j0; i0 += 128) units.push_back(make_int4(t, i0, j0, i0 + 64 < j0 ? 2 : 1));
            }
        } else if (chunk == MTE_TEA_MFMA) {
            for (int i0 = 0; i0 < n; i0 += 64, c++)
                for (int j0 = i0; j0 < n; j0 += 64) units.push_back(make_int4(t, i0, j0, c));
        } else {
            for (int j0 = 0; j0 < n; j0 += chunk, c++)
                for (int i0 = 0; i0 < n; i0 += 256) units.push_back(make_int4(t, i0, j0, c));
        }
        nchunks[t] = c;
        maxc = std::max(maxc, c);
    }
    if (chunk == MTE_TEA_SYMMETRIC || chunk == MTE_TEA_MFMA) {
        // longest first: off-diagonal blocks (64 steps of both sides), then
        // diagonal blocks (one-sided, one step per bead of the chunk)
        auto cost = [&](const int4 &u) {
            const int n = (int)(e->off[u.x + 1] - e->off[u.x]);
            return u.y == u.z ? std::min(64, n - u.z) : (chunk == MTE_TEA_SYMMETRIC ? 128 * u.w : 128);
        };
        std::stable_sort(units.begin(), units.end(),
                         [&](const int4 &a, const int4 &b) { return cost(a) > cost(b); });
    }
    if (chunk == MTE_TEA_MFMA) {
        // Workgroups go to the 8 XCDs in turn (workgroup k to XCD k mod 8),
        // each with its own L2: unit k is taken from trajectory t = k mod 8
        // (mod 8), so that every trajectory's moment tables are read through
        // one L2.  The units of each class keep the longest-first order.
        constexpr int NXCD = 8;
        std::vector<std::vector<int4>> cls(NXCD);
        for (const int4 &u : units) cls[u.x % NXCD].push_back(u);
        std::vector<size_t> next(NXCD, 0);
        units.clear();
        for (bool left = true; left;) {
            left = false;
            for (int x = 0; x < NXCD; x++)
                if (next[x] < cls[x].size()) {
                    units.push_back(cls[x][next[x]++]);
                    left = true;
                }
        }
        std::vector<int2> tiles;
        std::vector<int> tbase(e->ntraj);
        for (int t = 0; t < e->ntraj; t++) {
            tbase[t] = (int)tiles.size();
            const int n = (int)(e->off[t + 1] - e->off[t]);
            for (int p0 = 0; p0 < n; p0 += 64) tiles.push_back(make_int2(t, p0));
        }
        for (int4 &u : units)
            u = make_int4(tbase[u.x] + u.y / 64, tbase[u.x] + u.z / 64, u.y / 64, u.z / 64);
        hipError_t s;
        e->tea.ntiles = (int)tiles.size();
        if ((s = dalloc(e, &e->tea.tiles, tiles.size())) != hipSuccess) return s;
        if ((s = dalloc(e, &e->tea.ctr, tiles.size())) != hipSuccess) return s;
        if ((s = dalloc(e, &e->tea.rec, tiles.size() * 4 * 64)) != hipSuccess) return s;
        if ((s = dalloc(e, &e->tea.rbead, tiles.size() * 64)) != hipSuccess) return s;
        uint16_t *mom = nullptr;
        if ((s = dalloc(e, &mom, tiles.size() * 2 * 64 * 64)) != hipSuccess) return s;
        e->tea.mom = mom;
        if ((s = hipMemcpyAsync(e->tea.tiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice,
                                e->stream)) != hipSuccess)
            return s;
    }
    hipError_t s;
    if ((s = dalloc(e, &e->tea.units, units.size())) != hipSuccess) return s;
    if ((s = dalloc(e, &e->tea.nchunks, (size_t)e->ntraj)) != hipSuccess) return s;
    if ((s = dalloc(e, &e->tea.partial, (size_t)maxc * e->nb)) != hipSuccess) return s;
    if ((s = hipMemcpyAsync(e->tea.units, units.data(), sizeof(int4) * units.size(), hipMemcpyHostToDevice,
                            e->stream)) != hipSuccess)
        return s;
    if ((s = hipMemcpyAsync(e->tea.nchunks, nchunks.data(), sizeof(int) * nchunks.size(), hipMemcpyHostToDevice,
                            e->stream)) != hipSuccess)
        return s;
    e->tea.nunits = (int)units.size();
    e->tea.max_chunks = maxc;
    e->tea.unit_chunk = chunk;
    e->tune.tea_chunk = chunk;
    return hipStreamSynchronize(e->stream);
}

// Lane order of the resident kernel: per trajectory the mobile beads in index
// order, then the fixed and extra ones, so that the last (partly filled) wave
// of a pass holds beads with nothing to compute where the trajectory has
// enough of them.  Any order gives the same results (each bead's sums do not
// depend on the lane that computes them).
int mte::resident_mobile_max(const mte_engine *e, const uint8_t *extra) {
    int most = 0;
    for (int t = 0; t < e->ntraj; t++) {
        int k = 0;
        for (int64_t b = e->off[t]; b < e->off[t + 1]; b++) k += (e->fixed_h[b] || extra[b]) ? 0 : 1;
        most = std::max(most, k);
    }
    return most;
}

hipError_t mte::upload_resident_order(mte_engine *e, const uint8_t *extra) {
    std::vector<int> order(e->nb);
    for (int t = 0; t < e->ntraj; t++) {
        const int64_t b0 = e->off[t], n = e->off[t + 1] - e->off[t];
        int k = 0;
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i < n; i++) {
                const bool still = e->fixed_h[b0 + i] || extra[b0 + i];
                if (still == (pass == 1)) order[b0 + k++] = i;
            }
    }
    e->res_mobile_max = resident_mobile_max(e, extra);
    hipError_t s = hipMemcpyAsync(e->res_order, order.data(), sizeof(int) * e->nb, hipMemcpyHostToDevice, e->stream);
    if (s != hipSuccess) return s;
    return hipStreamSynchronize(e->stream);
}

// Work units of the one-pass epsilon kernel: every pair (I group, J group >=
// I group) of 1024-bead groups of a trajectory, and its partial buffer.  Left
// out (the two-sided kernel computes epsilon) when the buffer would exceed
// EPS_PART_MAX_BYTES.
hipError_t mte::tea_build_eps_units(mte_engine *e) {
    std::vector<int4> units;
    int ngrp = 0;
    for (int t = 0; t < e->ntraj; t++) {
        const int n = (int)(e->off[t + 1] - e->off[t]);
        const int g = (n + EPS_GROUP - 1) / EPS_GROUP;
        ngrp = std::max(ngrp, g);
        for (int i = 0; i < g; i++)
            for (int j = i; j < g; j++) units.push_back(make_int4(t, i, j, 0));
    }
    const size_t part = (size_t)2 * ngrp * 5 * e->nb;
    e->tune.tea_eps = MTE_TEA_EPS_TWOSIDED;
    if (units.empty() || part * sizeof(float) > EPS_PART_MAX_BYTES) return hipSuccess;
    hipError_t s;
    if ((s = dalloc(e, &e->tea.eps_units, units.size())) != hipSuccess) return s;
    if ((s = dalloc(e, &e->tea.eps_part, part)) != hipSuccess) return s;
    if ((s = hipMemcpyAsync(e->tea.eps_units, units.data(), sizeof(int4) * units.size(), hipMemcpyHostToDevice,
                            e->stream)) != hipSuccess)
        return s;
    e->tea.eps_nunits = (int)units.size();
    e->tea.eps_ngrp = ngrp;
    e->tune.tea_eps = MTE_TEA_EPS_ONEPASS;
    return hipStreamSynchronize(e->stream);
}

// src/ran2.h:18-56 with state in a struct; positive seed skips the table
// fill (shuffle table all zero, iy = 0, idum2 = 123456789).
namespace {
struct Ran2 {
    int32_t idum, idum2 = 123456789, iy = 0, iv[32] = {0};
    double next() {
        const int32_t IM1 = 2147483563, IM2 = 2147483399, IA1 = 40014, IA2 = 40692, IQ1 = 53668, IQ2 = 52774,
                      IR1 = 12211, IR2 = 3791, NTAB = 32, IMM1 = IM1 - 1, NDIV = 1 + IMM1 / NTAB;
        const double AM = 1.0 / IM1, RNMX = 1.0 - 1.2e-7;
        int32_t j, k;
        if (idum <= 0) {
            idum = (-idum < 1) ? 1 : -idum;
            idum2 = idum;
            for (j = NTAB + 7; j >= 0; j--) {
                k = idum / IQ1;
                idum = IA1 * (idum - k * IQ1) - k * IR1;
                if (idum < 0) idum += IM1;
                if (j < NTAB) iv[j] = idum;
            }
            iy = iv[0];
        }
        k = idum / IQ1;
        idum = IA1 * (idum - k * IQ1) - k * IR1;
        if (idum < 0) idum += IM1;
        k = idum2 / IQ2;
        idum2 = IA2 * (idum2 - k * IQ2) - k * IR2;
        if (idum2 < 0) idum2 += IM2;
        j = iy / NDIV;
        iy = iv[j] - idum2;
        iv[j] = idum;
        if (iy < 1) iy += IMM1;
        const double t = AM * iy;
        return t > RNMX ? RNMX : t;
    }
};
}  // namespace

// generateSeeds, src/HybridTaus.cu:32-48.
static void generate_seeds(int32_t seed, size_t nwords, uint32_t *out) {
    Ran2 r;
    r.idum = seed;
    for (size_t i = 0; i < nwords; i++) {
        uint32_t v;
        do { v = (uint32_t)(r.next() * 4294967295.0); } while (v < 128);
        out[i] = v;
    }
}

namespace mte {
// The list-capacity flags into the host's pinned, device-mapped words.
__global__ void k_publish_flags(const int *overflow, const int *frozen, int *out) {
    out[0] = *(volatile const int *)overflow;
    out[1] = *(volatile const int *)frozen;
}
}  // namespace mte

// Kernel timing with HIP events on the engine's stream.  Both events of a
// pair are taken before the start is recorded, so no host work sits
// between the kernel and its stop event; read events go back to a pool.
struct EvPair {
    hipEvent_t start = nullptr, stop = nullptr;
};

static hipEvent_t take_event(mte_engine *e) {
    if (!e->ev_pool.empty()) {
        const hipEvent_t ev = e->ev_pool.back();
        e->ev_pool.pop_back();
        return ev;
    }
    // Timing only: without the system-scope fence a marker writes back no
    // caches, so it neither costs the following work nor skews the times.
    hipEvent_t ev = nullptr;
    (void)hipEventCreateWithFlags(&ev, hipEventDisableSystemFence);
    return ev;
}

static void time_begin(mte_engine *e, const char *name, EvPair *ev) {
    if (!e->timing) return;
    ev->start = take_event(e);
    ev->stop = take_event(e);
    (void)hipEventRecord(ev->start, e->stream);
    (void)name;
}

static void time_end(mte_engine *e, const char *name, const EvPair &ev) {
    if (!e->timing) return;
    (void)hipEventRecord(ev.stop, e->stream);
    e->timers[name].pending.emplace_back(ev.start, ev.stop);
}

// Reads every pending pair; an event may end one pair and start the next
// (a rebuild and the resident launch after it), so each returns to the pool
// once.
static void flush_timers(mte_engine *e) {
    std::vector<hipEvent_t> done;
    for (auto &kv : e->timers) {
        for (auto &pr : kv.second.pending) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) kv.second.total_ms += ms;
            kv.second.launches++;
            done.push_back(pr.first);
            done.push_back(pr.second);
        }
        kv.second.pending.clear();
    }
    std::sort(done.begin(), done.end());
    done.erase(std::unique(done.begin(), done.end()), done.end());
    e->ev_pool.insert(e->ev_pool.end(), done.begin(), done.end());
}

extern "C" {

const char *mte_last_error(const mte_engine *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int mte_destroy(mte_engine *e) {
    if (!e) return MTE_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    flush_timers(e);
    for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
    if (e->flags_host) (void)hipHostFree(e->flags_host);
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    for (void *p : e->allocations) (void)hipFree(p);
    for (void *p : {(void *)e->mlp.x, (void *)e->mlp.t, (void *)e->mlp.hid, (void *)e->mlp.y, (void *)e->mlp.g4,
                    (void *)e->mlp.ga, (void *)e->mlp.gb, (void *)e->mlp.partial})
        if (p) (void)hipFree(p);
    if (e->noise) (void)hipFree(e->noise);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return MTE_OK;
}

int mte_create(const mte_params *params, const mte_topology *topo, const mte_options *opt, mte_engine **out) {
    if (!out || !params || !topo || !opt) {
        g_create_error = "mte_create: null argument";
        return MTE_ERR_ARG;
    }
    *out = nullptr;
    mte_engine *e = new mte_engine();
    auto bail = [&](int code) {
        g_create_error = e->err;
        mte_destroy(e);
        return code;
    };
    e->P = *params;
    e->O = *opt;
    if (e->O.max_lj <= 0) e->O.max_lj = 64;
    if (e->O.max_lj > 1023) { e->err = "mte_create: max_lj must be at most 1023"; return bail(MTE_ERR_ARG); }
    const int nt = topo->ntraj;
    if (nt <= 0 || !topo->traj_offsets) { e->err = "mte_create: ntraj must be positive"; return bail(MTE_ERR_ARG); }
    e->off.assign(topo->traj_offsets, topo->traj_offsets + nt + 1);
    if (e->off[0] != 0) { e->err = "mte_create: traj_offsets[0] must be 0"; return bail(MTE_ERR_ARG); }
    for (int t = 0; t < nt; t++)
        if (e->off[t + 1] <= e->off[t]) { e->err = "mte_create: empty or decreasing trajectory"; return bail(MTE_ERR_ARG); }
    if (e->off[nt] >= (int64_t(1) << 29)) { e->err = "mte_create: too many beads"; return bail(MTE_ERR_ARG); }
    if (e->P.ljpairsupdatefreq <= 0) { e->err = "mte_create: LJPairsUpdateFreq must be positive"; return bail(MTE_ERR_ARG); }
    if (e->P.tea_on && e->P.tea_epsilon_freq <= 0) { e->err = "mte_create: tea_epsilon_freq must be positive"; return bail(MTE_ERR_ARG); }
    const int nb = (int)e->off[nt];
    e->nb = nb;
    e->ntraj = nt;
    e->tune.stream_block = 256;
    e->tune.tea_chunk = 0;   // set when TEA's work units are built
    e->tune.cell_stencil = 1;
    e->device = opt->device;
    if (hipSetDevice(e->device) != hipSuccess) { e->err = "mte_create: hipSetDevice failed"; return bail(MTE_ERR_HIP); }
    if (hipDeviceGetAttribute(&e->num_cus, hipDeviceAttributeMultiprocessorCount, e->device) != hipSuccess ||
        e->num_cus < 1) {
        e->err = "mte_create: cannot query the compute-unit count";
        return bail(MTE_ERR_HIP);
    }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        e->err = "mte_create: stream creation failed";
        return bail(MTE_ERR_HIP);
    }
    if (noalloc_scratch_init(e->device, &e->noalloc) != hipSuccess) {
        e->err = "mte_create: hipGetSymbolAddress failed for the inference staging array";
        return bail(MTE_ERR_HIP);
    }

    // Lateral offsets p1, p2 (src/mt.h:44-50) in double, rounded once.
    const double c = std::cos(2.0 * M_PI / 13.0), s = std::sin(2.0 * M_PI / 13.0);
    e->L = LatOffsets{(float)(0.5 * 8.12 * (c - 1.0)), (float)(0.5 * 8.12 * s), (float)(-3.0 * 2.0 / 13.0),
                      (float)(0.5 * 8.12 * (c - 1.0)), (float)(-0.5 * 8.12 * s), (float)(3.0 * 2.0 / 13.0)};

    // Host-side topology translation: local reference encodings -> packed
    // global entries, k-major list slots.
    std::vector<int> tid(nb), tbase(nt), tsize(nt), harm(nb), harm_raw(nb), lon_n(nb, 0), lat_n(nb, 0);
    std::vector<int> lon((size_t)MAXLON * nb, 0), lat((size_t)MAXLAT * nb, 0);
    for (int t = 0; t < nt; t++) {
        tbase[t] = (int)e->off[t];
        tsize[t] = (int)(e->off[t + 1] - e->off[t]);
        for (int i = 0; i < tsize[t]; i++) tid[tbase[t] + i] = t;
    }
    for (int b = 0; b < nb; b++) {
        const int t = tid[b], n = tsize[t], base = tbase[t];
        const int hc = topo->harmonic_count ? topo->harmonic_count[b] : 0;
        harm_raw[b] = topo->harmonic ? topo->harmonic[b] : -1;
        if (hc > 1) { e->err = "mte_create: at most one harmonic partner per bead"; return bail(MTE_ERR_ARG); }
        if (hc == 1) {
            const int h = harm_raw[b];
            const int j = h < 0 ? -h : h;
            if (j >= n) { e->err = "mte_create: harmonic partner out of range"; return bail(MTE_ERR_ARG); }
            harm[b] = ((base + j) << 1) | (h < 0 ? 1 : 0);
        } else {
            harm[b] = -1;
        }
        const int nl = topo->longitudinal_count ? topo->longitudinal_count[b] : 0;
        const int nr = topo->lateral_count ? topo->lateral_count[b] : 0;
        if (nl < 0 || nl > MAXLON || nr < 0 || nr > MAXLAT) { e->err = "mte_create: list count out of range"; return bail(MTE_ERR_ARG); }
        for (int k = 0; k < nl; k++) {
            const int v = topo->longitudinal[(size_t)b * MAXLON + k];
            const int j = v < 0 ? -v : v;
            if (j >= n) { e->err = "mte_create: longitudinal entry out of range"; return bail(MTE_ERR_ARG); }
            lon[(size_t)k * nb + b] = ((base + j) << 1) | (v < 0 ? 1 : 0);
        }
        for (int k = 0; k < nr; k++) {
            const int v = topo->lateral[(size_t)b * MAXLAT + k];
            int j = v <= 0 ? -v : v;
            if (j == 999999) j = 0;
            if (j >= n) { e->err = "mte_create: lateral entry out of range"; return bail(MTE_ERR_ARG); }
            lat[(size_t)k * nb + b] = ((base + j) << 1) | (v <= 0 ? 1 : 0);
        }
        lon_n[b] = nl;
        lat_n[b] = nr;
    }
    std::vector<int2> tiles;
    for (int t = 0; t < nt; t++)
        for (int i0 = 0; i0 < tsize[t]; i0 += 256) tiles.push_back(make_int2(t, i0));
    e->ntiles = (int)tiles.size();

    DevTopo &T = e->T;
    T.nb = nb;
    T.ntraj = nt;
    T.max_lj = e->O.max_lj;
    int *p_tid, *p_tbase, *p_tsize, *p_harm, *p_harm_raw, *p_mon, *p_gtp, *p_ont;
    uint8_t *p_fixed, *p_extra;
    float *p_zs, *p_box;
#define ALLOC(ptr, n)                                                          \
    if (dalloc(e, &(ptr), (n)) != hipSuccess) {                                \
        e->err = "mte_create: hipMalloc failed for " #ptr;                     \
        return bail(MTE_ERR_HIP);                                              \
    }
    ALLOC(p_tid, nb) ALLOC(p_tbase, nt) ALLOC(p_tsize, nt) ALLOC(p_harm, nb) ALLOC(p_harm_raw, nb)
    ALLOC(T.lon_n, nb) ALLOC(T.lon, (size_t)MAXLON * nb) ALLOC(T.lat_n, nb) ALLOC(T.lat, (size_t)MAXLAT * nb)
    ALLOC(T.lj_n, nb) ALLOC(T.lj, (size_t)T.max_lj * nb) ALLOC(p_fixed, nb) ALLOC(p_extra, nb) ALLOC(p_mon, nb)
    ALLOC(p_gtp, nb) ALLOC(p_ont, nb) ALLOC(p_zs, nt) ALLOC(p_box, nt) ALLOC(T.overflow, 1) ALLOC(T.frozen, 1) ALLOC(e->tiles, tiles.size())
    ALLOC(e->rb_order, nb) ALLOC(e->res_order, nb) ALLOC(e->mtl_counts, nt) ALLOC(e->mtl_flags, nb)
    for (int k = 0; k < 2; k++) {
        ALLOC(e->S[k].xyz, nb) ALLOC(e->S[k].ang, nb) ALLOC(e->S[k].fn, nb) ALLOC(e->S[k].fp1, nb)
        ALLOC(e->S[k].fp2, nb)
    }
    ALLOC(e->seeds, (size_t)2 * nb) ALLOC(e->fxyz, nb) ALLOC(e->fang, nb) ALLOC(e->en_bead, (size_t)7 * nb)
    ALLOC(e->en_traj, (size_t)7 * nt)
    if (e->P.tea_on) {
        ALLOC(e->tea.ci, nb) ALLOC(e->tea.eps, nb) ALLOC(e->tea.beta, nt) ALLOC(e->tea.stats, 2 * nt)
        ALLOC(e->tea.status, 1) ALLOC(e->tea.R, nb)
        if (tea_build_units(e, TEA_CHUNK_BEADS) != hipSuccess || tea_build_eps_units(e) != hipSuccess) {
            e->err = "mte_create: TEA work units";
            return bail(MTE_ERR_HIP);
        }
    }
#undef ALLOC
    T.tid = p_tid; T.tbase = p_tbase; T.tsize = p_tsize; T.harm = p_harm; T.harm_raw = p_harm_raw;
    T.fixed = p_fixed; T.extra = p_extra; T.mon_type = p_mon; T.gtp = p_gtp; T.on_tub = p_ont; T.zs = p_zs; T.tbox = p_box;

    std::vector<uint8_t> fixed(nb, 0), extra(nb, 0);
    std::vector<int> mon(nb, 0), gtp(nb, 1), ont(nb, 0);
    std::vector<float> zs(nt, e->P.rep_h);
    for (int b = 0; b < nb; b++) {
        if (topo->fixed) fixed[b] = topo->fixed[b] ? 1 : 0;
        if (topo->extra) extra[b] = topo->extra[b] ? 1 : 0;
        if (topo->mon_type) mon[b] = topo->mon_type[b];
        if (topo->gtp) gtp[b] = topo->gtp[b];
        if (topo->on_tubule) ont[b] = topo->on_tubule[b];
    }
    if (topo->wall_z) std::copy(topo->wall_z, topo->wall_z + nt, zs.begin());
    std::vector<float> box(nt, 0.f);
    if (topo->box) std::copy(topo->box, topo->box + nt, box.begin());
    for (int t = 0; t < nt; t++) {
        if (!(box[t] >= 0.f)) { e->err = "mte_create: periodic box side must be >= 0"; return bail(MTE_ERR_ARG); }
        if (box[t] > 0.f) e->periodic = true;
    }
    std::vector<uint32_t> seeds((size_t)8 * nb);
    generate_seeds(opt->rseed, seeds.size(), seeds.data());

#define UP(dst, src, n)                                                                         \
    if (hipMemcpyAsync((void *)(dst), (src), (size_t)(n) * sizeof(*(src)), hipMemcpyHostToDevice, \
                       e->stream) != hipSuccess) {                                              \
        e->err = "mte_create: upload failed for " #dst;                                         \
        return bail(MTE_ERR_HIP);                                                               \
    }
    UP(p_tid, tid.data(), nb) UP(p_tbase, tbase.data(), nt) UP(p_tsize, tsize.data(), nt)
    UP(p_harm, harm.data(), nb) UP(p_harm_raw, harm_raw.data(), nb) UP(T.lon_n, lon_n.data(), nb)
    UP(T.lon, lon.data(), lon.size()) UP(T.lat_n, lat_n.data(), nb) UP(T.lat, lat.data(), lat.size())
    UP(p_fixed, fixed.data(), nb) UP(p_extra, extra.data(), nb) UP(p_mon, mon.data(), nb)
    UP(p_gtp, gtp.data(), nb) UP(p_ont, ont.data(), nb) UP(p_zs, zs.data(), nt)
    UP(p_box, box.data(), nt)
    UP(e->tiles, tiles.data(), tiles.size()) UP(e->seeds, seeds.data(), seeds.size())
    std::vector<int> order(nb);   // rebuild order: index order until a state arrives (mte_set_state)
    for (int t = 0; t < nt; t++)
        for (int i = 0; i < tsize[t]; i++) order[tbase[t] + i] = i;
    UP(e->rb_order, order.data(), nb)
#undef UP
    e->fixed_h = fixed;
    e->any_extra = std::any_of(extra.begin(), extra.end(), [](uint8_t v) { return v != 0; });
    if (upload_resident_order(e, extra.data()) != hipSuccess) {
        e->err = "mte_create: upload failed for the resident lane order";
        return bail(MTE_ERR_HIP);
    }
    if (hipMemsetAsync(T.lj_n, 0, sizeof(int) * nb, e->stream) != hipSuccess ||
        hipMemsetAsync(T.overflow, 0, sizeof(int), e->stream) != hipSuccess ||
        hipMemsetAsync(T.frozen, 0, sizeof(int), e->stream) != hipSuccess) {
        e->err = "mte_create: memset failed";
        return bail(MTE_ERR_HIP);
    }
    if (hipHostMalloc((void **)&e->flags_host, 2 * sizeof(int), hipHostMallocMapped | hipHostMallocCoherent) !=
            hipSuccess ||
        hipHostGetDevicePointer((void **)&e->flags_dev, e->flags_host, 0) != hipSuccess) {
        e->err = "mte_create: pinned allocation failed for the list-capacity flags";
        return bail(MTE_ERR_HIP);
    }
    e->flags_host[0] = e->flags_host[1] = 0;
    if (e->P.tea_on && hipMemsetAsync(e->tea.status, 0, sizeof(int), e->stream) != hipSuccess) {
        e->err = "mte_create: memset failed";
        return bail(MTE_ERR_HIP);
    }
    // Cell lists: selection and tables.
    for (int t = 0; t < nt; t++) e->nmax = std::max(e->nmax, tsize[t]);
    {
        const int nmax = e->nmax;
        const float radius = candidate_radius(e->P);
        bool cells_ok = radius > 0.f;
        for (int t = 0; t < nt; t++)
            if (box[t] > 0.f && box[t] < 3.f * radius) cells_ok = false;   // 27-cell search needs >= 3 cells per side
        if (opt->list_method == MTE_LISTS_CELLS && !cells_ok) {
            e->err = "mte_create: cell lists need every periodic box side >= 3 x the list radius";
            return bail(MTE_ERR_ARG);
        }
        e->use_cells = cells_ok && (opt->list_method == MTE_LISTS_CELLS ||
                                    (opt->list_method == MTE_LISTS_AUTO && nmax > 2048));
        if (e->use_cells || e->P.tea_on) {   // MLP-HI's neighbour search needs the tables too
            std::vector<int> toff(nt), tsz(nt);
            int total = 0;
            for (int t = 0; t < nt; t++) {
                int s = 64;
                while (s < 2 * tsize[t]) s <<= 1;
                toff[t] = total;
                tsz[t] = s;
                total += s;
            }
            CellTables &C = e->cells;
            int *p_toff, *p_tsz;
            if (dalloc(e, &p_toff, nt) != hipSuccess || dalloc(e, &p_tsz, nt) != hipSuccess ||
                dalloc(e, &C.count, total) != hipSuccess || dalloc(e, &C.rank, nb) != hipSuccess ||
                dalloc(e, &C.cell, nb) != hipSuccess ||
                dalloc(e, &C.sorted, nb) != hipSuccess || dalloc(e, &C.range, total) != hipSuccess ||
                dalloc(e, &C.spos, nb) != hipSuccess ||
                dalloc(e, &C.bsum, (total + CELL_SCAN_SPAN - 1) / CELL_SCAN_SPAN) != hipSuccess) {
                e->err = "mte_create: hipMalloc failed for cell tables";
                return bail(MTE_ERR_HIP);
            }
            C.toff = p_toff;
            C.tsize = p_tsz;
            C.total = total;
            if (hipMemsetAsync(C.count, 0, sizeof(int) * (size_t)total, e->stream) != hipSuccess ||
                hipMemcpyAsync(p_toff, toff.data(), sizeof(int) * nt, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
                hipMemcpyAsync(p_tsz, tsz.data(), sizeof(int) * nt, hipMemcpyHostToDevice, e->stream) != hipSuccess) {
                e->err = "mte_create: upload failed for cell tables";
                return bail(MTE_ERR_HIP);
            }
        }
    }
    if (e->P.tea_on) {
        MlpBuffers &M = e->mlp;
        if (dalloc(e, &M.w, MLP_NPARAM) != hipSuccess || dalloc(e, &M.wi, MLP_NPARAM) != hipSuccess ||
            dalloc(e, &M.grad, MLP_NPARAM) != hipSuccess || dalloc(e, &M.m1, MLP_NPARAM) != hipSuccess ||
            dalloc(e, &M.m2, MLP_NPARAM) != hipSuccess || dalloc(e, &M.feat, (size_t)MLP_IN * nb) != hipSuccess ||
            dalloc(e, &M.fblk, nb) != hipSuccess ||
            dalloc(e, &M.targ, (size_t)MLP_OUT * nb) != hipSuccess ||
            dalloc(e, &M.ystep, (size_t)MLP_OUT * nb) != hipSuccess || dalloc(e, &M.err, 2 * (size_t)nt) != hipSuccess ||
            dalloc(e, &M.err_part, 2 * (size_t)nt * MLP_ERR_PARTS) != hipSuccess ||
            dalloc(e, &M.nbad, 1) != hipSuccess || dalloc(e, &M.img, mlp_image_bytes() / sizeof(uint4)) != hipSuccess) {
            e->err = "mte_create: hipMalloc failed for MLP buffers";
            return bail(MTE_ERR_HIP);
        }
        if (hipMemsetAsync(M.nbad, 0, sizeof(int), e->stream) != hipSuccess) {
            e->err = "mte_create: memset failed";
            return bail(MTE_ERR_HIP);
        }
        e->mlp_last_err.assign(nt, 0.0);
        // Neighbour-search workgroups round robin over the trajectories: the
        // first blocks of every trajectory (a tube's crowded beads, which
        // come first in the usual input, run longest) are dispatched first,
        // and with a multiple of 8 trajectories workgroup k keeps to the
        // XCD of trajectory k mod 8.
        std::vector<int2> bl;
        int most = 0;
        for (int t = 0; t < nt; t++) most = std::max(most, (tsize[t] + MLP_SEARCH_BLOCK - 1) / MLP_SEARCH_BLOCK);
        for (int k = 0; k < most; k++)
            for (int t = 0; t < nt; t++)
                if (k * MLP_SEARCH_BLOCK < tsize[t]) bl.push_back(make_int2(t, k));
        M.nblocks = (int)bl.size();
        if (dalloc(e, &M.blocks, bl.size()) != hipSuccess ||
            hipMemcpyAsync(M.blocks, bl.data(), sizeof(int2) * bl.size(), hipMemcpyHostToDevice, e->stream) !=
                hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) {
            e->err = "mte_create: upload failed for the neighbour-search blocks";
            return bail(MTE_ERR_HIP);
        }
    }
    {   // list rebuilds with one workgroup per trajectory when the largest one fits in LDS
        int nmax = 0, lds_limit = 0;
        for (int t = 0; t < nt; t++) nmax = std::max(nmax, tsize[t]);
        if (hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) != hipSuccess ||
            !rebuild_traj_plan(nmax, lds_limit, &e->rb_traj_threads, &e->rb_traj_lds))
            e->rb_traj_threads = 0;
        // forces with the trajectory staged in LDS when it fits
        e->forces_lds = forces_lds_bytes(nmax) <= (size_t)lds_limit ? forces_lds_bytes(nmax) : 0;
        e->tune.forces = e->forces_lds > 0 ? MTE_FORCES_LDS : MTE_FORCES_GLOBAL;
    }
    ResidentPlan plan;
    const bool fits = resident_fits(e, &plan);
    if (opt->mode == MTE_MODE_RESIDENT && !fits) {
        e->err = "mte_create: resident mode requested but the batch does not fit (TEA on, or a trajectory too large)";
        return bail(MTE_ERR_ARG);
    }
    e->mode = (opt->mode == MTE_MODE_STREAMED || !fits) ? MTE_MODE_STREAMED : MTE_MODE_RESIDENT;
    e->rplan = plan;
    e->tune.resident_per_thread = fits ? plan.per_thread : 0;
    e->tune.mode = e->mode;
    if (hipStreamSynchronize(e->stream) != hipSuccess) { e->err = "mte_create: sync failed"; return bail(MTE_ERR_HIP); }
    *out = e;
    return MTE_OK;
}

int mte_num_beads(const mte_engine *e, int64_t *nbeads, int32_t *ntraj) {
    if (!e) return MTE_ERR_ARG;
    if (nbeads) *nbeads = e->nb;
    if (ntraj) *ntraj = e->ntraj;
    return MTE_OK;
}

int mte_get_mode(const mte_engine *e, int32_t *mode) {
    if (!e || !mode) return MTE_ERR_ARG;
    *mode = e->mode;
    return MTE_OK;
}

int mte_set_state(mte_engine *e, const mte_coord *r) {
    if (!e || !r) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    std::vector<float4> xyz(e->nb), ang(e->nb);
    for (int b = 0; b < e->nb; b++) {
        xyz[b] = make_float4(r[b].x, r[b].y, r[b].z, 0.f);
        ang[b] = make_float4(r[b].fi, r[b].theta, r[b].psi, 0.f);
    }
    HIPCHK(hipMemcpyAsync(e->S[e->cur].xyz, xyz.data(), sizeof(float4) * e->nb, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->S[e->cur].ang, ang.data(), sizeof(float4) * e->nb, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_frames(e, e->cur));
    HIPCHK(launch_rebuild_order(e));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_get_state(mte_engine *e, mte_coord *r) {
    if (!e || !r) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    std::vector<float4> xyz(e->nb), ang(e->nb);
    HIPCHK(hipMemcpyAsync(xyz.data(), e->S[e->cur].xyz, sizeof(float4) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(ang.data(), e->S[e->cur].ang, sizeof(float4) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int b = 0; b < e->nb; b++) {
        r[b].x = xyz[b].x; r[b].y = xyz[b].y; r[b].z = xyz[b].z;
        r[b].fi = ang[b].x; r[b].theta = ang[b].y; r[b].psi = ang[b].z; r[b].w = 0.f;
    }
    return MTE_OK;
}

int mte_set_flags(mte_engine *e, const uint8_t *extra, const int32_t *gtp, const int32_t *on_tubule) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if (extra) {
        std::vector<uint8_t> x(extra, extra + e->nb);
        for (auto &v : x) v = v ? 1 : 0;
        // The resident plan for the new flags is decided before anything
        // changes: a one-bead-per-thread plan over a trajectory longer than
        // the workgroup needs its beads past the workgroup to stay immobile.
        const int mobile_old = e->res_mobile_max;
        e->res_mobile_max = resident_mobile_max(e, x.data());
        ResidentPlan plan = e->rplan;
        bool plan_ok = true;
        if (e->rplan.per_thread == 1 && !resident_plan_for(e, 1, &plan)) plan_ok = resident_fits(e, &plan);
        if (!plan_ok && e->mode == MTE_MODE_RESIDENT) {
            e->res_mobile_max = mobile_old;   // engine unchanged
            return fail(e, MTE_ERR_STATE, "mte_set_flags: no resident plan fits the new flags");
        }
        HIPCHK(hipMemcpyAsync(e->T.extra, x.data(), e->nb, hipMemcpyHostToDevice, e->stream));
        e->any_extra = std::any_of(x.begin(), x.end(), [](uint8_t v) { return v != 0; });
        HIPCHK(upload_resident_order(e, x.data()));
        HIPCHK(hipStreamSynchronize(e->stream));
        if (plan_ok) {
            e->rplan = plan;
            if (e->tune.resident_per_thread) e->tune.resident_per_thread = plan.per_thread;
        } else {   // streamed engine: no resident plan fits these flags
            e->rplan = ResidentPlan{};
            e->tune.resident_per_thread = 0;
        }
    }
    if (gtp) HIPCHK(hipMemcpyAsync(e->T.gtp, gtp, sizeof(int) * e->nb, hipMemcpyHostToDevice, e->stream));
    if (on_tubule) HIPCHK(hipMemcpyAsync(e->T.on_tub, on_tubule, sizeof(int) * e->nb, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_mt_length(mte_engine *e, float rad_hi, float rad_lo, float cos_min, int32_t apply, int32_t *counts,
                  int32_t *on_tubule) {
    if (!e || !counts) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(launch_mt_length(e, rad_hi, rad_lo, cos_min, apply ? 1 : 0));
    HIPCHK(hipMemcpyAsync(counts, e->mtl_counts, sizeof(int) * e->ntraj, hipMemcpyDeviceToHost, e->stream));
    if (on_tubule)
        HIPCHK(hipMemcpyAsync(on_tubule, e->mtl_flags, sizeof(int) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_set_step(mte_engine *e, int64_t step) {
    if (!e || step < 0) return MTE_ERR_ARG;
    e->step = step;
    return MTE_OK;
}

int mte_get_step(const mte_engine *e, int64_t *step) {
    if (!e || !step) return MTE_ERR_ARG;
    *step = e->step;
    return MTE_OK;
}

// The list-capacity flag of the rebuilds since the last check.  After an
// overflow the resident kernel leaves the state and RNG streams at the
// overflowing rebuild's step (T.frozen counts the steps it did not take), so
// the step counter and the injected-noise cursor go back by that many; the
// next run rebuilds the lists again (and overflows again unless the state
// changed).
//
// The flags reach the host through k_publish_flags, enqueued after the work
// that may set them and read once the stream has drained: one small kernel
// in place of two device-to-host copies and a synchronisation of their own.
static hipError_t publish_flags(mte_engine *e) {
    hipLaunchKernelGGL(mte::k_publish_flags, dim3(1), dim3(1), 0, e->stream, e->T.overflow, e->T.frozen,
                       e->flags_dev);
    return hipGetLastError();
}

// The published flags, after a stream synchronisation that followed
// publish_flags.
static int eval_overflow(mte_engine *e) {
    const int ov = ((volatile int *)e->flags_host)[0], frozen = ((volatile int *)e->flags_host)[1];
    if (!ov) return MTE_OK;
    HIPCHK(hipMemsetAsync(e->T.overflow, 0, sizeof(int), e->stream));
    HIPCHK(hipMemsetAsync(e->T.frozen, 0, sizeof(int), e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->step -= frozen;
    if (e->noise) e->noise_cursor -= frozen;
    e->lists_step = -1;
    e->lists_valid = false;
    char buf[200];
    std::snprintf(buf, sizeof buf,
                  "neighbour list capacity exceeded (longitudinal 8, lateral 16 or LJ max_lj) by the rebuild at "
                  "step %lld; the state is that step's",
                  (long long)e->step);
    return fail(e, MTE_ERR_OVERFLOW, buf);
}

static int check_overflow(mte_engine *e) {
    HIPCHK(publish_flags(e));
    HIPCHK(hipStreamSynchronize(e->stream));
    return eval_overflow(e);
}

// `end`, when given, receives the rebuild's stop event (null while timing is
// off), for a launch that follows at once to start its timing from.
static int rebuild_now(mte_engine *e, hipEvent_t *end = nullptr) {
    EvPair ev;
    time_begin(e, "rebuild", &ev);
    HIPCHK(launch_rebuild(e));
    time_end(e, "rebuild", ev);
    if (end) *end = ev.stop;
    e->lists_step = e->step;
    e->lists_valid = true;
    return MTE_OK;
}

static bool rebuild_due(const mte_engine *e) {
    return (e->step % e->P.ljpairsupdatefreq) == 0 && e->lists_step != e->step;
}

int mte_rebuild_lists(mte_engine *e) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    int rc = rebuild_now(e);
    if (rc) return rc;
    return check_overflow(e);
}

static int check_tea(mte_engine *e) {
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, e->tea.status, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    if (st) {
        (void)hipMemsetAsync(e->tea.status, 0, sizeof(int), e->stream);
        char buf[160];
        std::snprintf(buf, sizeof buf, "HI tensor check failed for trajectory %d at step %lld (epsilon/beta out of range)",
                      st - 1, (long long)e->step);
        return fail(e, MTE_ERR_TEA, buf);
    }
    return MTE_OK;
}

static bool mlp_active(const mte_engine *e) { return e->hi_mode == MTE_HI_MLP && !e->mlp_fallback; }

// Enqueue `len` streamed steps starting at global step s0 (TEA refreshes and
// MLP checks at their periods; a list rebuild due at s0 has run already),
// flipping e->cur per step.  The same sequence is launched directly or
// recorded into a graph.
static hipError_t enqueue_streamed(mte_engine *e, int64_t s0, int64_t len, int64_t noise_cursor) {
    hipError_t st = hipSuccess;
    for (int64_t k = 0; k < len && st == hipSuccess; k++) {
        const int64_t step = s0 + k;
        const float *nz = e->noise ? e->noise + (size_t)6 * e->nb * (noise_cursor + k) : nullptr;
        EvPair ev;
        if (mlp_active(e)) {
            // forces, TEA normalisation refresh, features (+ exact labels on
            // check steps), fused prediction + integration
            const bool check = step % e->mlp_check_every == 0;
            st = launch_forces(e);
            if (st == hipSuccess && step % e->P.tea_epsilon_freq == 0) st = launch_tea_eps(e);
            if (st == hipSuccess) st = launch_mlp_cells(e);
            if (st != hipSuccess) break;
            time_begin(e, "mlp_local", &ev);
            st = launch_mlp_local(e, e->mlp.feat, check ? e->mlp.targ : nullptr, false, e->mlp.fblk);
            time_end(e, "mlp_local", ev);
            if (st != hipSuccess) break;
            time_begin(e, "mlp_step", &ev);
            st = launch_mlp_step(e, nz, check ? e->mlp.ystep : nullptr);
            time_end(e, "mlp_step", ev);
            if (st == hipSuccess && check) st = launch_mlp_error(e, e->mlp.ystep, e->mlp.targ);
        } else if (e->P.tea_on) {
            st = launch_forces_tea(e, nz);
            if (st == hipSuccess && step % e->P.tea_epsilon_freq == 0) st = launch_tea_eps(e);
            if (st != hipSuccess) break;
            time_begin(e, "tea_pairs", &ev);
            st = launch_tea_pairs(e);
            time_end(e, "tea_pairs", ev);
            if (st == hipSuccess) st = launch_tea_finish(e, nz);
        } else {
            time_begin(e, "step", &ev);
            st = launch_step_plain(e, nz);
            time_end(e, "step", ev);
        }
    }
    return st;
}

// Steps from `step` to the next list rebuild, TEA refresh or MLP check
// (exclusive), so that a segment only carries events at its first step.
static int64_t steps_to_next_event(const mte_engine *e, int64_t step) {
    const int64_t f = e->P.ljpairsupdatefreq;
    int64_t n = f - (step % f);
    if (e->P.tea_on) n = std::min<int64_t>(n, e->P.tea_epsilon_freq - (step % e->P.tea_epsilon_freq));
    if (mlp_active(e)) n = std::min<int64_t>(n, e->mlp_check_every - (step % e->mlp_check_every));
    return n;
}

// Sampled-error judge of MLP-HI: relative error of the predicted A - I
// against the exact labels per trajectory; above the tolerance the run
// continues with TEA (the fallback), whose epsilon/beta are refreshed first.
static int check_mlp(mte_engine *e) {
    std::vector<double> se(2 * e->ntraj);
    int bad = 0;
    HIPCHK(hipMemcpyAsync(se.data(), e->mlp.err, sizeof(double) * se.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(&bad, e->mlp.nbad, sizeof(int), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->mlp_bad = bad;
    double worst = 0.0;
    for (int t = 0; t < e->ntraj; t++) {
        const double r = se[2 * t + 1] > 0 ? std::sqrt(se[2 * t] / se[2 * t + 1]) : 0.0;
        e->mlp_last_err[t] = r;
        worst = std::max(worst, r);
    }
    if (worst > e->mlp_tol || bad > 0) {
        e->mlp_fallback = true;
        HIPCHK(launch_tea_eps(e));
        int rc = check_tea(e);
        if (rc) return rc;
        e->tea_valid = true;
    }
    return MTE_OK;
}

int mte_run(mte_engine *e, int64_t nsteps) {
    if (!e || nsteps < 0) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if (e->noise && e->noise_cursor + nsteps > e->noise_steps)
        return fail(e, MTE_ERR_NOISE, "injected noise exhausted");
    int64_t done = 0;
    const int64_t freq = e->P.ljpairsupdatefreq;
    const bool need_lists = e->P.lj_on || e->P.is_assembly;
    while (done < nsteps) {
        if (e->mode == MTE_MODE_RESIDENT) {
            // One launch covers every step up to the next rebuild boundary.
            const int64_t to_boundary = freq - (e->step % freq);
            const int64_t chunk = std::min<int64_t>(nsteps - done, to_boundary);
            // lists due at this boundary are rebuilt by the streamed kernel first (same stream)
            hipEvent_t mark = nullptr;   // the rebuild's stop marker, which also starts the step's timing
            if (need_lists && rebuild_due(e)) {
                const int rc = rebuild_now(e, &mark);
                if (rc) return rc;
            }
            const float *nz = e->noise ? e->noise + (size_t)6 * e->nb * e->noise_cursor : nullptr;
            EvPair ev;
            if (mark) {
                ev.start = mark;
                ev.stop = take_event(e);
            } else {
                time_begin(e, "step", &ev);
            }
            HIPCHK(launch_resident(e, (int)chunk, nz));
            time_end(e, "step", ev);
            if (e->noise) e->noise_cursor += chunk;
            e->step += chunk;
            done += chunk;
            continue;
        }
        const int64_t seg = std::min<int64_t>(nsteps - done, steps_to_next_event(e, e->step));
        const bool mlp_now = mlp_active(e);
        const bool mlp_check = mlp_now && e->step % e->mlp_check_every == 0;
        const bool tea_refresh = e->P.tea_on && e->step % e->P.tea_epsilon_freq == 0;
        if (e->P.tea_on && !e->tea_valid && !tea_refresh) {
            // TEA's epsilon/beta must exist before the first TEA or MLP-HI step
            HIPCHK(launch_tea_eps(e));
            int rc = check_tea(e);
            if (rc) return rc;
            e->tea_valid = true;
        }
        if (need_lists && rebuild_due(e)) {
            // Lists are rebuilt outside the captured steps, and their
            // capacity flag read, so that an overflow is reported before any
            // step runs on truncated lists.
            int rc = rebuild_now(e);
            if (!rc) rc = check_overflow(e);
            if (rc) return rc;
        }
        if (e->use_graphs && !e->noise && !e->timing) {
            const int64_t hi_phase = (e->P.tea_on ? e->step % e->P.tea_epsilon_freq : 0) +
                                     (mlp_now ? 1000003 * (1 + e->step % e->mlp_check_every) : 0);
            const auto key = std::make_tuple(hi_phase, seg, e->cur);
            auto it = e->graphs.find(key);
            const int cur0 = e->cur;
            if (it == e->graphs.end()) {
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                const hipError_t st = enqueue_streamed(e, e->step, seg, 0);
                const hipError_t st2 = hipStreamEndCapture(e->stream, &g);
                e->cur = cur0;
                HIPCHK(st);
                HIPCHK(st2);
                hipGraphExec_t ex;
                const hipError_t st3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                HIPCHK(st3);
                it = e->graphs.emplace(key, ex).first;
            }
            HIPCHK(hipGraphLaunch(it->second, e->stream));
            e->cur = cur0 ^ (int)(seg & 1);
        } else {
            HIPCHK(enqueue_streamed(e, e->step, seg, e->noise_cursor));
        }
        if (tea_refresh) {
            int rc = check_tea(e);
            if (rc) return rc;
            e->tea_valid = true;
        }
        if (mlp_check) {
            int rc = check_mlp(e);
            if (rc) return rc;
        }

        if (e->noise) e->noise_cursor += seg;
        e->step += seg;
        done += seg;
    }
    if (need_lists) HIPCHK(publish_flags(e));
    HIPCHK(hipStreamSynchronize(e->stream));
    flush_timers(e);
    if (need_lists) return eval_overflow(e);
    return MTE_OK;
}

int mte_compute_forces(mte_engine *e, mte_coord *f) {
    if (!e || !f) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if ((e->P.lj_on || e->P.is_assembly) && rebuild_due(e)) {
        int rc = rebuild_now(e);
        if (rc) return rc;
    }
    HIPCHK(launch_forces(e));
    std::vector<float4> fx(e->nb), fa(e->nb);
    HIPCHK(hipMemcpyAsync(fx.data(), e->fxyz, sizeof(float4) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(fa.data(), e->fang, sizeof(float4) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int b = 0; b < e->nb; b++) {
        f[b].x = fx[b].x; f[b].y = fx[b].y; f[b].z = fx[b].z;
        f[b].fi = fa[b].x; f[b].theta = fa[b].y; f[b].psi = fa[b].z; f[b].w = 0.f;
    }
    return check_overflow(e);
}

int mte_compute_energies(mte_engine *e, mte_energies *per_bead, double *per_traj) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if ((e->P.lj_on || e->P.is_assembly) && rebuild_due(e)) {
        int rc = rebuild_now(e);
        if (rc) return rc;
    }
    HIPCHK(launch_energies(e));
    if (per_bead)
        HIPCHK(hipMemcpyAsync(per_bead, e->en_bead, sizeof(double) * 7 * e->nb, hipMemcpyDeviceToHost, e->stream));
    if (per_traj)
        HIPCHK(hipMemcpyAsync(per_traj, e->en_traj, sizeof(double) * 7 * e->ntraj, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return check_overflow(e);
}

int mte_get_lists(mte_engine *e, int32_t *lon_count, int32_t *lon, int32_t *lat_count, int32_t *lat,
                  int32_t *lj_count, int32_t *lj) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    const int nb = e->nb, ml = e->T.max_lj;
    std::vector<int> ln(nb), l((size_t)MAXLON * nb), rn(nb), r((size_t)MAXLAT * nb), jn(nb), j((size_t)ml * nb);
    HIPCHK(hipMemcpyAsync(ln.data(), e->T.lon_n, sizeof(int) * nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(l.data(), e->T.lon, sizeof(int) * l.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(rn.data(), e->T.lat_n, sizeof(int) * nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(r.data(), e->T.lat, sizeof(int) * r.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(jn.data(), e->T.lj_n, sizeof(int) * nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipMemcpyAsync(j.data(), e->T.lj, sizeof(int) * j.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<int> tbase(nb);
    for (int t = 0; t < e->ntraj; t++)
        for (int64_t b = e->off[t]; b < e->off[t + 1]; b++) tbase[b] = (int)e->off[t];
    for (int b = 0; b < nb; b++) {
        if (lon_count) lon_count[b] = ln[b];
        if (lat_count) lat_count[b] = rn[b];
        if (lj_count) lj_count[b] = jn[b];
        if (lon)
            for (int k = 0; k < MAXLON; k++) {
                const int v = l[(size_t)k * nb + b];
                const int jj = (v >> 1) - tbase[b];
                lon[(size_t)b * MAXLON + k] = k < ln[b] ? ((v & 1) ? -jj : jj) : 0;
            }
        if (lat)
            for (int k = 0; k < MAXLAT; k++) {
                const int v = r[(size_t)k * nb + b];
                const int jj = (v >> 1) - tbase[b];
                int enc = (v & 1) ? (jj != 0 ? -jj : -999999) : (jj != 0 ? jj : 999999);
                lat[(size_t)b * MAXLAT + k] = k < rn[b] ? enc : 0;
            }
        if (lj)
            for (int k = 0; k < ml; k++) lj[(size_t)b * ml + k] = k < jn[b] ? j[(size_t)k * nb + b] - tbase[b] : 0;
    }
    return MTE_OK;
}

int mte_set_noise(mte_engine *e, const float *noise, int64_t nsteps) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if (e->noise) {
        HIPCHK(hipStreamSynchronize(e->stream));
        HIPCHK(hipFree(e->noise));
        e->noise = nullptr;
    }
    e->noise_steps = 0;
    e->noise_cursor = 0;
    if (!noise || nsteps <= 0) return MTE_OK;
    const size_t bytes = sizeof(float) * 6 * (size_t)e->nb * nsteps;
    HIPCHK(hipMalloc(&e->noise, bytes));
    HIPCHK(hipMemcpyAsync(e->noise, noise, bytes, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->noise_steps = nsteps;
    return MTE_OK;
}

int mte_get_rng(mte_engine *e, uint32_t *seeds) {
    if (!e || !seeds) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(seeds, e->seeds, sizeof(uint4) * 2 * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_set_rng(mte_engine *e, const uint32_t *seeds) {
    if (!e || !seeds) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(e->seeds, seeds, sizeof(uint4) * 2 * e->nb, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_tea_stats(mte_engine *e, double *eps, double *beta) {
    if (!e) return MTE_ERR_ARG;
    if (!e->P.tea_on || !e->tea_valid) return fail(e, MTE_ERR_STATE, "no TEA refresh has run");
    HIPCHK(hipSetDevice(e->device));
    std::vector<double> st(2 * e->ntraj);
    HIPCHK(hipMemcpyAsync(st.data(), e->tea.stats, sizeof(double) * st.size(), hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int t = 0; t < e->ntraj; t++) {
        if (eps) eps[t] = st[2 * t];
        if (beta) beta[t] = st[2 * t + 1];
    }
    return MTE_OK;
}

int mte_tea_ci(mte_engine *e, float *ci) {
    if (!e || !ci) return MTE_ERR_ARG;
    if (!e->P.tea_on || !e->tea_valid) return fail(e, MTE_ERR_STATE, "no TEA refresh has run");
    HIPCHK(hipSetDevice(e->device));
    std::vector<float4> c(e->nb);
    HIPCHK(hipMemcpyAsync(c.data(), e->tea.ci, sizeof(float4) * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int b = 0; b < e->nb; b++) {
        ci[3 * b] = c[b].x;
        ci[3 * b + 1] = c[b].y;
        ci[3 * b + 2] = c[b].z;
    }
    return MTE_OK;
}

// ------------------------------------------------------------------ MLP-HI
static int need_mlp(mte_engine *e) {
    if (!e) return MTE_ERR_ARG;
    if (!e->P.tea_on) return fail(e, MTE_ERR_STATE, "MLP-HI needs TEA parameters (tea_on, tea_a)");
    std::vector<float> box(e->ntraj);
    if (hipMemcpy(box.data(), e->T.tbox, sizeof(float) * e->ntraj, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(e, MTE_ERR_HIP, "need_mlp: box download failed");
    for (float L : box)
        if (L > 0.f && L < 36.f) return fail(e, MTE_ERR_ARG, "MLP-HI needs periodic boxes of at least 36 nm");
    return MTE_OK;
}

static int mlp_reserve(mte_engine *e, int n) {
    MlpBuffers &M = e->mlp;
    if (n <= M.cap) return MTE_OK;
    for (void *p : {(void *)M.x, (void *)M.t, (void *)M.hid, (void *)M.y, (void *)M.g4, (void *)M.ga, (void *)M.gb,
                    (void *)M.partial})
        if (p) {
            (void)hipStreamSynchronize(e->stream);
            (void)hipFree(p);
        }
    M.cap = 0;
    const size_t nn = (size_t)n;
    if (hipMalloc(&M.x, sizeof(float) * nn * MLP_IN) != hipSuccess ||
        hipMalloc(&M.t, sizeof(float) * nn * MLP_OUT) != hipSuccess ||
        hipMalloc(&M.hid, sizeof(float) * 3 * nn * MLP_H) != hipSuccess ||
        hipMalloc(&M.y, sizeof(float) * nn * MLP_OUT) != hipSuccess ||
        hipMalloc(&M.g4, sizeof(float) * nn * MLP_OUT) != hipSuccess ||
        hipMalloc(&M.ga, sizeof(float) * nn * MLP_H) != hipSuccess ||
        hipMalloc(&M.gb, sizeof(float) * nn * MLP_H) != hipSuccess ||
        hipMalloc(&M.partial, sizeof(double) * ((nn * MLP_OUT + 255) / 256 + 1)) != hipSuccess)
        return fail(e, MTE_ERR_HIP, "MLP work space allocation failed");
    M.cap = n;
    return MTE_OK;
}

static double mlp_loss(mte_engine *e, int n) {
    const int nblk = (n * MLP_OUT + 255) / 256;
    std::vector<double> part(nblk);
    if (hipMemcpyAsync(part.data(), e->mlp.partial, sizeof(double) * nblk, hipMemcpyDeviceToHost, e->stream) !=
            hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return -1.0;
    double s = 0;
    for (double v : part) s += v;
    return s / ((double)n * MLP_OUT);
}

int mte_mlp_set_weights(mte_engine *e, const float *w, int64_t n) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!w || n != MLP_NPARAM) return fail(e, MTE_ERR_ARG, "mte_mlp_set_weights: expected MTE_MLP_NPARAM weights");
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipMemcpyAsync(e->mlp.w, w, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_mlp_round(e));
    HIPCHK(hipMemsetAsync(e->mlp.m1, 0, sizeof(float) * MLP_NPARAM, e->stream));
    HIPCHK(hipMemsetAsync(e->mlp.m2, 0, sizeof(float) * MLP_NPARAM, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    e->mlp.adam_t = 0;
    e->mlp.ready = true;
    return MTE_OK;
}

int mte_mlp_get_weights(mte_engine *e, float *w, int64_t n) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!w || n != MLP_NPARAM) return fail(e, MTE_ERR_ARG, "mte_mlp_get_weights: expected MTE_MLP_NPARAM weights");
    HIPCHK(hipMemcpyAsync(w, e->mlp.w, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_mlp_local(mte_engine *e, float *features, float *labels) {
    int rc = need_mlp(e);
    if (rc) return rc;
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(launch_mlp_local(e, e->mlp.feat, e->mlp.targ));
    if (features)
        HIPCHK(hipMemcpyAsync(features, e->mlp.feat, sizeof(float) * MLP_IN * e->nb, hipMemcpyDeviceToHost, e->stream));
    if (labels)
        HIPCHK(hipMemcpyAsync(labels, e->mlp.targ, sizeof(float) * MLP_OUT * e->nb, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_mlp_forward(mte_engine *e, const float *features, int64_t n, float *out) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!features || !out || n <= 0) return fail(e, MTE_ERR_ARG, "mte_mlp_forward: bad arguments");
    HIPCHK(hipSetDevice(e->device));
    rc = mlp_reserve(e, (int)n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(e->mlp.x, features, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_mlp_forward(e, e->mlp.x, (int)n, e->mlp.y));
    HIPCHK(hipMemcpyAsync(out, e->mlp.y, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_mlp_forward_noalloc(mte_engine *e, const float *features, int64_t n, float *out) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!features || !out || n <= 0) return fail(e, MTE_ERR_ARG, "mte_mlp_forward_noalloc: bad arguments");
    if (n > MLP_NOALLOC_ROWS)
        return fail(e, MTE_ERR_ARG, "mte_mlp_forward_noalloc: at most MTE_MLP_NOALLOC_ROWS (8192) rows per call");
    HIPCHK(hipSetDevice(e->device));
    // every row the kernel reads or writes is set by this call, so nothing
    // in the shared array needs clearing
    float *x = e->noalloc, *y = e->noalloc + (size_t)MLP_NOALLOC_ROWS * MLP_IN;
    HIPCHK(hipMemcpyAsync(x, features, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_mlp_forward(e, x, (int)n, y));
    HIPCHK(hipMemcpyAsync(out, y, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    return MTE_OK;
}

int mte_mlp_loss_grad(mte_engine *e, const float *features, const float *labels, int64_t n, double *loss,
                      float *grad) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!features || !labels || n <= 0) return fail(e, MTE_ERR_ARG, "mte_mlp_loss_grad: bad arguments");
    HIPCHK(hipSetDevice(e->device));
    rc = mlp_reserve(e, (int)n);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(e->mlp.x, features, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemcpyAsync(e->mlp.t, labels, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice, e->stream));
    HIPCHK(launch_mlp_grad(e, e->mlp.x, e->mlp.t, (int)n));
    if (grad) HIPCHK(hipMemcpyAsync(grad, e->mlp.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, e->stream));
    const double l = mlp_loss(e, (int)n);
    if (loss) *loss = l;
    return MTE_OK;
}

int mte_mlp_train(mte_engine *e, const float *features, const float *labels, int64_t n, int32_t iters,
                  int32_t batch, float lr, double *losses) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (!features || !labels || n <= 0 || iters < 0 || batch <= 0 || batch > n)
        return fail(e, MTE_ERR_ARG, "mte_mlp_train: bad arguments");
    if (!e->mlp.ready) return fail(e, MTE_ERR_STATE, "mte_mlp_train: set initial weights first");
    HIPCHK(hipSetDevice(e->device));
    // the whole training set stays on the device; batches are consecutive slices
    float *X = nullptr, *Y = nullptr;
    HIPCHK(hipMalloc(&X, sizeof(float) * MLP_IN * n));
    if (hipMalloc(&Y, sizeof(float) * MLP_OUT * n) != hipSuccess) {
        (void)hipFree(X);
        return fail(e, MTE_ERR_HIP, "mte_mlp_train: allocation failed");
    }
    rc = mlp_reserve(e, batch);
    int out = rc;
    if (!rc) {
        if (hipMemcpyAsync(X, features, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, e->stream) != hipSuccess ||
            hipMemcpyAsync(Y, labels, sizeof(float) * MLP_OUT * n, hipMemcpyHostToDevice, e->stream) != hipSuccess)
            out = fail(e, MTE_ERR_HIP, "mte_mlp_train: upload failed");
        int64_t pos = 0;
        for (int it = 0; it < iters && !out; it++) {
            if (pos + batch > n) pos = 0;
            hipError_t s = launch_mlp_grad(e, X + pos * MLP_IN, Y + pos * MLP_OUT, batch);
            if (s == hipSuccess) s = launch_mlp_adam(e, lr, ++e->mlp.adam_t);
            if (s != hipSuccess) out = fail(e, MTE_ERR_HIP, std::string("mte_mlp_train: ") + hipGetErrorString(s));
            if (losses && !out) losses[it] = mlp_loss(e, batch);
            pos += batch;
        }
        if (!out && launch_mlp_pack(e) != hipSuccess) out = fail(e, MTE_ERR_HIP, "mte_mlp_train: weight image failed");
        (void)hipStreamSynchronize(e->stream);
    }
    (void)hipFree(X);
    (void)hipFree(Y);
    return out;
}

int mte_set_hi_mode(mte_engine *e, int32_t mode, int32_t check_every, float tolerance) {
    if (!e) return MTE_ERR_ARG;
    if (mode != MTE_HI_DEFAULT && mode != MTE_HI_MLP) return fail(e, MTE_ERR_ARG, "mte_set_hi_mode: unknown mode");
    if (mode == MTE_HI_MLP) {
        int rc = need_mlp(e);
        if (rc) return rc;
        if (!e->mlp.ready) return fail(e, MTE_ERR_STATE, "mte_set_hi_mode: MLP weights not set");
        if (e->mode != MTE_MODE_STREAMED) return fail(e, MTE_ERR_STATE, "MLP-HI runs in streamed mode");
        if (check_every <= 0 || !(tolerance > 0.f)) return fail(e, MTE_ERR_ARG, "mte_set_hi_mode: bad check settings");
    }
    e->hi_mode = mode;
    e->mlp_check_every = check_every > 0 ? check_every : 100;
    e->mlp_tol = tolerance > 0.f ? tolerance : 0.02f;
    e->mlp_fallback = false;
    e->tea_valid = false;
    return MTE_OK;
}

int mte_mlp_status(mte_engine *e, double *last_error, int32_t *fallback, int64_t *bad_beads) {
    int rc = need_mlp(e);
    if (rc) return rc;
    if (last_error)
        for (int t = 0; t < e->ntraj; t++) last_error[t] = e->mlp_last_err[t];
    if (fallback) *fallback = e->mlp_fallback ? 1 : 0;
    if (bad_beads) *bad_beads = e->mlp_bad;
    return MTE_OK;
}

int mte_set_timing(mte_engine *e, int32_t on) {
    if (!e) return MTE_ERR_ARG;
    e->timing = on != 0;
    return MTE_OK;
}

int mte_reset_timing(mte_engine *e) {
    if (!e) return MTE_ERR_ARG;
    HIPCHK(hipStreamSynchronize(e->stream));
    flush_timers(e);
    e->timers.clear();
    return MTE_OK;
}

int mte_kernel_time(mte_engine *e, const char *name, double *total_ms, int64_t *launches) {
    if (!e || !name) return MTE_ERR_ARG;
    HIPCHK(hipStreamSynchronize(e->stream));
    flush_timers(e);
    auto it = e->timers.find(name);
    if (total_ms) *total_ms = it == e->timers.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == e->timers.end() ? 0 : it->second.launches;
    return MTE_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ tuning
static void clear_graphs(mte_engine *e) {
    for (auto &kv : e->graphs) (void)hipGraphExecDestroy(kv.second);
    e->graphs.clear();
}

int mte_get_tuning(const mte_engine *e, mte_tuning *t) {
    if (!e || !t) return MTE_ERR_ARG;
    *t = e->tune;
    return MTE_OK;
}

int mte_set_tuning(mte_engine *e, const mte_tuning *t) {
    if (!e || !t) return MTE_ERR_ARG;
    HIPCHK(hipSetDevice(e->device));
    if (t->mode) {
        ResidentPlan plan;
        if (t->mode == MTE_MODE_RESIDENT) {
            if (!resident_fits(e, &plan))
                return fail(e, MTE_ERR_ARG, "mte_set_tuning: the batch does not fit the resident kernel");
            if (e->mode != MTE_MODE_RESIDENT) {
                e->rplan = plan;
                e->tune.resident_per_thread = plan.per_thread;
            }
        } else if (t->mode != MTE_MODE_STREAMED) {
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: mode must be MTE_MODE_RESIDENT or MTE_MODE_STREAMED");
        }
        if (t->mode != e->mode) {
            // both paths rebuild their lists at the next rebuild step; switch there
            if ((e->P.lj_on || e->P.is_assembly) && e->step % e->P.ljpairsupdatefreq != 0)
                return fail(e, MTE_ERR_STATE, "mte_set_tuning: switch modes at a list rebuild step");
            e->mode = t->mode;
            e->lists_step = -1;
            e->lists_valid = false;
        }
        e->tune.mode = e->mode;
    }
    if (t->resident_per_thread) {
        // a plan for the resident kernel; kept but unused while the engine runs streamed
        ResidentPlan plan;
        if (!resident_plan_for(e, t->resident_per_thread, &plan))
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: resident_per_thread not usable for this batch");
        if (e->mode == MTE_MODE_RESIDENT) e->rplan = plan;
        e->tune.resident_per_thread = t->resident_per_thread;
    }
    if (t->stream_block) {
        if (t->stream_block != 64 && t->stream_block != 128 && t->stream_block != 256)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: stream_block must be 64, 128 or 256");
        e->tune.stream_block = t->stream_block;
    }
    if (t->tea_chunk && t->tea_chunk != e->tune.tea_chunk) {
        if (t->tea_chunk != 64 && t->tea_chunk != 128 && t->tea_chunk != 256 && t->tea_chunk != MTE_TEA_SYMMETRIC &&
            t->tea_chunk != MTE_TEA_MFMA)
            return fail(e, MTE_ERR_ARG,
                        "mte_set_tuning: tea_chunk must be 64, 128, 256, MTE_TEA_SYMMETRIC or MTE_TEA_MFMA");
        if (!e->P.tea_on) return fail(e, MTE_ERR_ARG, "mte_set_tuning: tea_chunk needs TEA");
        if (t->tea_chunk == MTE_TEA_MFMA && e->periodic)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: MTE_TEA_MFMA needs open boundaries");
        if (t->tea_chunk == MTE_TEA_MFMA && e->nmax > REBUILD_ORDER_MAX)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: MTE_TEA_MFMA needs trajectories of at most 2048 beads");
        if (t->tea_chunk != e->tea.unit_chunk) HIPCHK(tea_build_units(e, t->tea_chunk));
    }
    if (t->tea_eps) {
        if (t->tea_eps != MTE_TEA_EPS_ONEPASS && t->tea_eps != MTE_TEA_EPS_TWOSIDED)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: tea_eps must be MTE_TEA_EPS_ONEPASS or MTE_TEA_EPS_TWOSIDED");
        if (!e->P.tea_on) return fail(e, MTE_ERR_ARG, "mte_set_tuning: tea_eps needs TEA");
        if (t->tea_eps == MTE_TEA_EPS_ONEPASS && !e->tea.eps_units)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: the one-pass epsilon buffer exceeds its 6 GiB cap here");
        e->tune.tea_eps = t->tea_eps;
    }
    if (t->cell_stencil) {
        if (t->cell_stencil != 1 && t->cell_stencil != 2)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: cell_stencil must be 1 or 2");
        e->tune.cell_stencil = t->cell_stencil;
    }
    if (t->forces) {
        if (t->forces != MTE_FORCES_LDS && t->forces != MTE_FORCES_GLOBAL)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: forces must be MTE_FORCES_LDS or MTE_FORCES_GLOBAL");
        if (t->forces == MTE_FORCES_LDS && e->forces_lds == 0)
            return fail(e, MTE_ERR_ARG, "mte_set_tuning: MTE_FORCES_LDS needs every trajectory to fit in LDS");
        e->tune.forces = t->forces;
    }
    clear_graphs(e);   // captured launches hold the old shapes
    return MTE_OK;
}

int mte_autotune(mte_engine *e, int32_t trial_steps, mte_tuning *chosen, char *report, int64_t report_len) {
    if (!e || trial_steps <= 0) return MTE_ERR_ARG;
    if (e->noise) return fail(e, MTE_ERR_STATE, "mte_autotune: injected noise pending");
    const bool need_lists = e->P.lj_on || e->P.is_assembly;
    if ((need_lists && e->step % e->P.ljpairsupdatefreq != 0) ||
        (e->P.tea_on && e->step % e->P.tea_epsilon_freq != 0))
        return fail(e, MTE_ERR_STATE, "mte_autotune: call where lists and epsilon are due (e.g. step 0)");
    HIPCHK(hipSetDevice(e->device));
    std::vector<mte_coord> r(e->nb);
    std::vector<uint32_t> rng((size_t)8 * e->nb);
    int rc;
    if ((rc = mte_get_state(e, r.data())) || (rc = mte_get_rng(e, rng.data()))) return rc;
    const int64_t step0 = e->step;
    const bool fb0 = e->mlp_fallback;
    const bool timing0 = e->timing;
    e->timing = false;
    std::string log;
    auto restore = [&]() -> int {
        int q;
        if ((q = mte_set_state(e, r.data())) || (q = mte_set_rng(e, rng.data()))) return q;
        e->step = step0;
        e->lists_step = -1;
        e->lists_valid = false;
        e->tea_valid = false;
        e->mlp_fallback = fb0;
        return MTE_OK;
    };
    // one warm-up run (graph capture, first-touch) and one timed run per candidate
    auto trial_with = [&](const mte_tuning &t, const char *what, int value, double *ms) -> int {
        int q;
        if ((q = restore()) || (q = mte_set_tuning(e, &t)) || (q = mte_run(e, trial_steps)) || (q = restore()))
            return q;
        const auto t0 = std::chrono::steady_clock::now();
        if ((q = mte_run(e, trial_steps))) return q;
        *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / trial_steps;
        char line[160];
        std::snprintf(line, sizeof line, "%s %d: %.4f ms/step\n", what, value, *ms);
        log += line;
        return MTE_OK;
    };
    auto pick = [](mte_tuning &t, const char *what, int value) {
        if (!strcmp(what, "stream_block")) t.stream_block = value;
        else if (!strcmp(what, "tea_chunk")) t.tea_chunk = value;
        else if (!strcmp(what, "forces")) t.forces = value;
        else t.cell_stencil = value;
    };
    auto trial = [&](const char *what, int value, double *ms) -> int {
        mte_tuning t{};
        pick(t, what, value);
        return trial_with(t, what, value, ms);
    };
    auto family = [&](const char *what, std::initializer_list<int> values, int *field) -> int {
        int best = *field;
        double best_ms = 1e300;
        for (int v : values) {
            double ms = 0;
            const int q = trial(what, v, &ms);
            if (q) return q;
            if (ms < best_ms) {
                best_ms = ms;
                best = v;
            }
        }
        // mte_set_tuning applies the choice (and rebuilds what depends on it)
        (void)field;
        mte_tuning t{};
        pick(t, what, best);
        return mte_set_tuning(e, &t);
    };
    // plain dynamics: every resident plan against every streamed workgroup size
    ResidentPlan plan;
    const bool can_resident = resident_fits(e, &plan);
    if (!e->P.tea_on) {
        mte_tuning best{};
        double best_ms = 1e300;
        auto consider = [&](const mte_tuning &t, const char *what, int value) -> int {
            double ms = 0;
            int q = trial_with(t, what, value, &ms);
            if (q) return q;
            if (ms < best_ms) {
                best_ms = ms;
                best = t;
            }
            return MTE_OK;
        };
        if (can_resident)
            for (int pt = 1; pt <= 3; pt++)
                if (resident_plan_for(e, pt, &plan)) {
                    mte_tuning t{};
                    t.mode = MTE_MODE_RESIDENT;
                    t.resident_per_thread = pt;
                    if ((rc = consider(t, "resident_per_thread", pt))) return rc;
                }
        for (int bs : {64, 128, 256}) {
            mte_tuning t{};
            t.mode = MTE_MODE_STREAMED;
            t.stream_block = bs;
            if ((rc = consider(t, "stream_block", bs))) return rc;
        }
        if ((rc = restore()) || (rc = mte_set_tuning(e, &best))) return rc;
    } else {
        if (e->forces_lds > 0 && (rc = family("forces", {MTE_FORCES_LDS, MTE_FORCES_GLOBAL}, &e->tune.forces)))
            return rc;
        if ((rc = family("stream_block", {64, 128, 256}, &e->tune.stream_block))) return rc;
        if (e->periodic || e->nmax > REBUILD_ORDER_MAX) {
            if ((rc = family("tea_chunk", {64, 128, 256, MTE_TEA_SYMMETRIC}, &e->tune.tea_chunk))) return rc;
        } else if ((rc = family("tea_chunk", {64, 128, 256, MTE_TEA_SYMMETRIC, MTE_TEA_MFMA}, &e->tune.tea_chunk))) {
            return rc;
        }
    }
    if (e->mode == MTE_MODE_STREAMED && e->use_cells && need_lists &&
        (rc = family("cell_stencil", {1, 2}, &e->tune.cell_stencil)))
        return rc;
    if ((rc = restore())) return rc;
    e->timing = timing0;
    if (chosen) *chosen = e->tune;
    if (report && report_len > 0) {
        const size_t n = std::min<size_t>((size_t)report_len - 1, log.size());
        std::memcpy(report, log.data(), n);
        report[n] = '\0';
    }
    return MTE_OK;
}
