// updater.cpp -- see updater.hpp.
#include "updater.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <exception>
#include <thread>

namespace mth {

namespace {

constexpr float R_THRES = R_MON * 8;   // src/mt.h:30
constexpr float ANG_THRES = 1.0f;      // src/mt.h:29
constexpr int PF_NUMBER = 13;

clock_t g_t0;
long long g_rand_calls = 0;   // glibc rand() draws since srand(rseed): replayed on resume

// The reference's host randomness (glibc rand() after srand(rseed)), counted.
int host_rand() {
    g_rand_calls++;
    return std::rand();
}

// Checkpoint of a run at a step where nothing is pending: the engine state,
// both HybridTaus streams, the updater's flags and counts, and the number of
// rand() draws so far.  Host byte order; version in the magic.
constexpr char CKPT_MAGIC[8] = {'M', 'T', 'E', 'C', 'K', 'P', 'T', '1'};

struct Checkpoint {
    long long step = 0, rand_calls = 0;
    std::vector<mte_coord> coords;
    std::vector<uint32_t> rng;   // 8 words per bead: the xyz stream, then the angular stream
    std::vector<uint8_t> extra;
    std::vector<int32_t> gtp, on_cur, on_prev, mt_len;
};

template <class T> void put(FILE *f, const std::vector<T> &v) {
    if (std::fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) throw IoError("checkpoint: short write");
}
template <class T> void get(FILE *f, std::vector<T> &v, size_t n) {
    v.resize(n);
    if (std::fread(v.data(), sizeof(T), n, f) != n) throw IoError("checkpoint: file too short");
}

void write_checkpoint(const std::string &path, const Checkpoint &c, int ntr, int ntot) {
    const std::string tmp = path + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) throw IoError("cannot write checkpoint " + tmp);
    const int32_t dims[2] = {ntr, ntot};
    const int64_t counters[2] = {c.step, c.rand_calls};
    bool ok = std::fwrite(CKPT_MAGIC, 1, 8, f) == 8 && std::fwrite(dims, sizeof dims, 1, f) == 1 &&
              std::fwrite(counters, sizeof counters, 1, f) == 1;
    if (ok) {
        put(f, c.coords);
        put(f, c.rng);
        put(f, c.extra);
        put(f, c.gtp);
        put(f, c.on_cur);
        put(f, c.on_prev);
        put(f, c.mt_len);
    }
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) throw IoError("cannot write checkpoint " + path);
}

Checkpoint read_checkpoint(const std::string &path, int ntr, int ntot) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw IoError("cannot read checkpoint " + path);
    char magic[8];
    int32_t dims[2];
    int64_t counters[2];
    Checkpoint c;
    try {
        if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, CKPT_MAGIC, 8) != 0)
            throw IoError("checkpoint " + path + ": not a checkpoint of this format");
        if (std::fread(dims, sizeof dims, 1, f) != 1 || std::fread(counters, sizeof counters, 1, f) != 1)
            throw IoError("checkpoint " + path + ": file too short");
        if (dims[0] != ntr || dims[1] != ntot)
            throw IoError("checkpoint " + path + ": written for " + std::to_string(dims[0]) + " x " +
                          std::to_string(dims[1]) + " beads, this run has " + std::to_string(ntr) + " x " +
                          std::to_string(ntot));
        const size_t nb = (size_t)ntr * ntot;
        c.step = counters[0];
        c.rand_calls = counters[1];
        get(f, c.coords, nb);
        get(f, c.rng, 8 * nb);
        get(f, c.extra, nb);
        get(f, c.gtp, nb);
        get(f, c.on_cur, nb);
        get(f, c.on_prev, nb);
        get(f, c.mt_len, ntr);
    } catch (...) {
        std::fclose(f);
        throw;
    }
    std::fclose(f);
    return c;
}

void print_formatted_time(float t) {
    const int days = (int)(t / (3600.0f * 24.0f));
    const int hours = (int)(t / 3600.0f - days * 24.0f);
    const int minutes = (int)(t / 60.0f - hours * 60.0f - days * 24.0f * 60.0f);
    const int seconds = (int)(t - hours * 3600.0f - days * 24.0f * 3600.0f - minutes * 60.0f);
    std::printf("%dd %dh %dm %ds", days, hours, minutes, seconds);
}

// printTime / printEstimatedTimeleft, src/timer.cpp (processor time, as there)
void print_time(long long step, long long steps) {
    const float t = (float)(std::clock() - g_t0) / (float)CLOCKS_PER_SEC;
    std::printf("Computation time: ");
    print_formatted_time(t);
    if (step != 0) std::printf(" (~%f steps/sec)\n", (float)step / t);
    else std::printf("\n");
    const float frac = (float)step / (float)steps;
    if (frac != 0.0f) {
        std::printf("Estimated time left: ");
        print_formatted_time(t * (1.0f / frac - 1.0f));
        std::printf(" (%3.1f%% completed)\n", frac * 100.0f);
    }
}

struct Engine {
    mte_engine *e = nullptr;
    ~Engine() {
        if (e) mte_destroy(e);
    }
};

// One DCD frame at a time on a worker thread: submit() waits for the
// previous frame, copies the state and starts the write; finish() waits for
// the last one.  An error of a write is rethrown by the next submit() or
// finish(); the destructor only joins.
class FrameWriter {
public:
    explicit FrameWriter(const Updater &up) : up_(up) {}
    ~FrameWriter() {
        if (th_.joinable()) th_.join();
    }
    void submit(const std::vector<mte_coord> &r) {
        finish();
        frame_ = r;
        th_ = std::thread([this]() {
            try {
                up_.save_dcd(frame_);
            } catch (...) {
                err_ = std::current_exception();
            }
        });
    }
    void finish() {
        if (th_.joinable()) th_.join();
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    const Updater &up_;
    std::vector<mte_coord> frame_;
    std::thread th_;
    std::exception_ptr err_;
};

int fail(mte_engine *e, const char *what, int rc) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, e ? mte_last_error(e) : "no engine");
    return rc;
}

}  // namespace

Updater::Updater(RunSetup &s, const RunOptions &o)
    : s_(s), o_(o), dcd_(s.ntot, (int)(s.stride != 0 ? s.steps / s.stride : 0), 1, s.dt, (int)s.stride) {}

void Updater::output_energies(const std::vector<double> &e) const {
    for (int t = 0; t < s_.ntr; t++) {
        const double *u = &e[7 * (size_t)t];
        if (!std::isfinite(u[0]) || !std::isfinite(u[1]) || !std::isfinite(u[2]) || !std::isfinite(u[6]))
            std::printf("Some energy in %d trajectory is NaN. NOT Exit program\n", t);
        std::printf("Energies[%d]:\t%f\t%f\t%f\t%f\t%f\t%f\t%f\n", t, u[0], u[1], u[2], u[3], u[4], u[5], u[6]);
    }
}

void Updater::output_forces(const std::vector<mte_coord> &f, const std::vector<mte_coord> &r) const {
    // OutputSumForce (float accumulation, :45-57) then OutputForces (:59-72)
    for (int t = 0; t < s_.ntr; t++) {
        float sx = 0, sy = 0, sz = 0;
        for (int i = t * s_.ntot; i < (t + 1) * s_.ntot; i++) {
            sx += f[i].x;
            sy += f[i].y;
            sz += f[i].z;
        }
        std::printf("SF for %d traj %.16f\n", t, std::sqrt(sx * sx + sy * sy + sz * sz));
    }
    for (size_t i = 0; i < f.size(); i++) {
        std::printf("Force[%d].theta = %f, fi = %f, psi = %f\n", (int)i, f[i].theta, f[i].fi, f[i].psi);
        std::printf("Angle[%d].theta = %f, fi = %f, psi = %f\n", (int)i, r[i].theta, r[i].fi, r[i].psi);
    }
}

int Updater::mt_length(mte_engine *e, long long step, std::vector<int> &mt_len) {
    if (step == 0) {
        FILE *f = std::fopen(o_.mt_len_file.c_str(), "w");
        if (f) std::fclose(f);
    }
    // the classification runs on the device state; the flags reach the
    // barrier there except at step 0, as the reference uploads them only
    // for step != 0 (src/compute_cuda.cu:1184-1193)
    std::vector<int32_t> counts(s_.ntr);
    const int rc = mte_mt_length(e, R_MT + R_THRES, 1.0f, std::cos(ANG_THRES), step != 0 ? 1 : 0, counts.data(),
                                 s_.on_tubule_cur.data());
    if (rc) return rc;
    for (int t = 0; t < s_.ntr; t++) {
        mt_len[t] = counts[t];
        if (!o_.quiet) std::printf("tubule[%d]: %d\n", t, counts[t]);
    }
    FILE *f = std::fopen(o_.mt_len_file.c_str(), "a");
    if (!f) throw IoError(o_.mt_len_file + ": cannot append");
    std::fprintf(f, "%lld\t", step);
    for (int t = 0; t < s_.ntr; t++) std::fprintf(f, "%f\t", 2 * R_MON * (float)mt_len[t] / PF_NUMBER);
    std::fprintf(f, "\n");
    std::fclose(f);
    return MTE_OK;
}

int Updater::change_conc(const std::vector<int> &delta, const std::vector<int> &mt_len, std::vector<mte_coord> &r) {
    (void)delta;   // passed by the reference's caller, unused there too
    int flag = 0;
    const int n = s_.ntot;
    for (int t = 0; t < s_.ntr; t++) {
        int num_extra = 0;
        for (int i = 0; i < n; i++) num_extra += s_.extra[i + (size_t)t * n] ? 1 : 0;
        const float vol = float(3.14 * s_.rep_r * s_.rep_r * s_.zs[t]);
        int nfree = (n - mt_len[t] - num_extra) / 2;
        while (1.0e7 * nfree / 6.0 < s_.conc * vol) {
            for (int i = 0; i < n; i += 2) {
                const size_t p = i + (size_t)t * n;
                if (s_.extra[p] && s_.mon_type[i] == 0) {
                    s_.extra[p] = 0;
                    s_.extra[p + 1] = 0;
                    float x, y;
                    for (;;) {
                        x = s_.rep_r - 2 * (host_rand() % int(s_.rep_r));
                        y = s_.rep_r - 2 * (host_rand() % int(s_.rep_r));
                        if (x * x + y * y <= s_.rep_r * s_.rep_r) {
                            if (!o_.quiet)
                                std::printf("New x,y coordinates for extra particle: %f  %f index: %d\n", x, y, (int)p);
                            num_extra -= 2;
                            break;
                        }
                    }
                    const float z = s_.zs[t] + s_.rep_leftborder + 3 * 2 * R_MON;
                    r[p].x = x;
                    r[p].y = y;
                    r[p].z = z;
                    r[p + 1].x = x;
                    r[p + 1].y = y;
                    r[p + 1].z = z + 2 * R_MON;
                    flag++;
                    break;
                }
            }
            nfree += 1;
            if (num_extra == 0) {
                if (!o_.quiet) std::printf("No more extra particles for trajectory[%d]!\n", t);
                break;
            }
        }
        if (!o_.quiet)
            std::printf("Concentration for tajectory[%d]: %f [muMole / L],\t %f [Dimers / nm^3],\t %d [Dimers / "
                        "Volume],\t  Volume: %f [nm^3]\n",
                        t, 1.0e7 * nfree / (6.0 * vol), nfree / vol, nfree, vol);
    }
    return flag;
}

void Updater::hydrolyse() {
    const int n = s_.ntot;
    for (int i = 0; i < n; i += 2)
        for (int t = 0; t < s_.ntr; t++) {
            const size_t p = i + (size_t)t * n;
            if (s_.gtp[p] == 1 && !s_.extra[p] && s_.on_tubule_cur[p] * s_.on_tubule_prev[p] == 1) {
                const double prob = host_rand() / (double)RAND_MAX;
                if (prob < 0.02) {
                    s_.gtp[p] = 0;
                    s_.gtp[p + 1] = 0;
                    if (!o_.quiet)
                        std::printf("*** Hydrolysis occured to dimer # %d trajectory #%d ***\n", i / 2, t);
                }
            }
        }
    for (int i = 0; i < n; i += 2)
        for (int t = 0; t < s_.ntr; t++) {
            const size_t p = i + (size_t)t * n;
            if (s_.gtp[p] == 0 && !s_.extra[p] && s_.on_tubule_cur[p] == 0 && s_.on_tubule_prev[p] == 0) {
                s_.gtp[p] = 1;
                s_.gtp[p + 1] = 1;
                if (!o_.quiet) std::printf("*** Transition to GTP occured to dimer # %d trajectory #%d ***\n", i / 2, t);
            }
        }
}

void Updater::create_dcds() const {
    for (int t = 0; t < s_.ntr; t++) {
        dcd_.create(s_.dcd_xyz[t]);
        dcd_.create(s_.dcd_ang[t]);
    }
}

void Updater::save_dcd(const std::vector<mte_coord> &r) const {
    const int n = s_.ntot;
    std::vector<float> X(n), Y(n), Z(n);
    for (int t = 0; t < s_.ntr; t++) {
        for (int i = 0; i < n; i++) {
            const mte_coord &c = r[(size_t)t * n + i];
            X[i] = c.x;
            Y[i] = c.y;
            Z[i] = c.z;
        }
        dcd_.append(s_.dcd_xyz[t], X.data(), Y.data(), Z.data());
        for (int i = 0; i < n; i++) {
            const mte_coord &c = r[(size_t)t * n + i];
            X[i] = c.fi;
            Y[i] = c.psi;
            Z[i] = c.theta;
        }
        dcd_.append(s_.dcd_ang[t], X.data(), Y.data(), Z.data());
    }
}

void Updater::append_hydrolysis_pdb(long long step, const std::vector<mte_coord> &r) const {
    // appendCoordPDB, src/preparator.cpp:621-647, into appendPDB's layout
    if (step == 0) {
        FILE *f = std::fopen(o_.hydrolysis_pdb.c_str(), "w");
        if (f) std::fclose(f);
    }
    FILE *f = std::fopen(o_.hydrolysis_pdb.c_str(), "a");
    if (!f) throw IoError(o_.hydrolysis_pdb + ": cannot append");
    const size_t total = (size_t)s_.ntot * s_.ntr;
    for (size_t k = 0; k < total; k++) {
        const PdbAtom &a = s_.atoms[k % s_.ntot];
        std::fprintf(f, total < 100000 ? "ATOM  %5d %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                                       : "ATOM  %5x %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n",
                     (int)(k + 1), a.name.c_str(), a.altLoc, a.resName.c_str(), a.chain, a.resid, (double)r[k].x,
                     (double)r[k].y, (double)r[k].z, (double)s_.gtp[k], (double)(k / s_.ntot));
    }
    std::fprintf(f, "END\n");
    std::fclose(f);
}

void Updater::save_result_pdb(const std::vector<mte_coord> &r) const {
    // saveCoordPDB, src/preparator.cpp:568-583: trajectory 0
    std::vector<PdbAtom> a = s_.atoms;
    for (int i = 0; i < s_.ntot; i++) {
        a[i].x = r[i].x;
        a[i].y = r[i].y;
        a[i].z = r[i].z;
    }
    write_pdb(o_.result_xyz, a);
    for (int i = 0; i < s_.ntot; i++) {
        a[i].x = r[i].fi;
        a[i].y = r[i].psi;
        a[i].z = r[i].theta;
    }
    write_pdb(o_.result_ang, a);
}

int run(RunSetup &s, const RunOptions &o) {
    if (s.stride == 0) {
        std::fprintf(stderr, "stride 0: the reference divides by it\n");
        return MTE_ERR_ARG;
    }
    std::srand((unsigned)s.rseed);   // main.cpp:67
    g_rand_calls = 0;
    g_t0 = std::clock();
    if (o.checkpoint_every < 0 ||
        (o.checkpoint_every > 0 && (o.checkpoint_every % std::llabs(s.stride) != 0 ||
                                    o.checkpoint_every % s.ljpairsupdatefreq != 0 ||
                                    (s.tea_on && o.checkpoint_every % s.tea_epsilon_freq != 0) ||
                                    (s.hydrolysis && s.hydrostep > 0 && o.checkpoint_every % s.hydrostep != 0)))) {
        std::fprintf(stderr, "engine_checkpoint %lld: must be a multiple of stride, LJPairsUpdateFreq, "
                             "tea_epsilon_freq (TEA) and the hydrolysis step\n", o.checkpoint_every);
        return MTE_ERR_ARG;
    }
    const int n = s.ntot, nt = s.ntr;
    const size_t nb = (size_t)n * nt;
    wrap_angles(s.coords);           // initIntegration wraps the host copy, then uploads it

    // per-bead topology for the CSR batch (identical trajectories)
    std::vector<int64_t> off(nt + 1);
    for (int t = 0; t <= nt; t++) off[t] = (int64_t)t * n;
    std::vector<int32_t> hc(nb), h(nb), mon(nb);
    std::vector<uint8_t> fixed(nb);
    for (size_t p = 0; p < nb; p++) {
        hc[p] = s.harmonic_count[p % n];
        h[p] = s.harmonic[p % n];
        mon[p] = s.mon_type[p % n];
        fixed[p] = s.fixed[p % n];
    }
    mte_topology topo{};
    topo.ntraj = nt;
    topo.traj_offsets = off.data();
    topo.harmonic_count = hc.data();
    topo.harmonic = h.data();
    topo.longitudinal_count = s.lon_count.data();
    topo.longitudinal = s.lon.data();
    topo.lateral_count = s.lat_count.data();
    topo.lateral = s.lat.data();
    topo.fixed = fixed.data();
    topo.extra = s.extra.data();
    topo.mon_type = mon.data();
    topo.gtp = s.gtp.data();
    topo.on_tubule = s.on_tubule_cur.data();
    topo.wall_z = s.zs.data();
    topo.box = nullptr;
    mte_options opt{};
    opt.max_lj = o.max_lj;
    opt.mode = o.mode;
    opt.rseed = s.rseed;
    opt.device = s.device;
    opt.list_method = MTE_LISTS_AUTO;
    const mte_params P = engine_params(s);
    Engine E;
    int rc = mte_create(&P, &topo, &opt, &E.e);
    if (rc) return fail(E.e, "mte_create", rc);
    if ((rc = mte_set_state(E.e, s.coords.data()))) return fail(E.e, "mte_set_state", rc);
    if (o.autotune_steps > 0) {
        mte_tuning t{};
        char report[4096];
        if ((rc = mte_autotune(E.e, o.autotune_steps, &t, report, sizeof report))) return fail(E.e, "mte_autotune", rc);
        if (!o.quiet) std::printf("Autotune:\n%s", report);
    }

    long long step = 0;
    std::vector<int> mt_len(nt, 0), mt_len_prev(nt, 0);
    if (!o.resume_file.empty()) {
        // continue a checkpointed run: the engine and the updater exactly as
        // they were when the checkpoint's step was about to run
        const Checkpoint c = read_checkpoint(o.resume_file, nt, n);
        s.coords = c.coords;
        s.extra = c.extra;
        s.gtp = c.gtp;
        s.on_tubule_cur = c.on_cur;
        s.on_tubule_prev = c.on_prev;
        mt_len.assign(c.mt_len.begin(), c.mt_len.end());
        for (long long k = 0; k < c.rand_calls; k++) host_rand();
        if ((rc = mte_set_state(E.e, s.coords.data())) || (rc = mte_set_rng(E.e, c.rng.data())) ||
            (rc = mte_set_flags(E.e, s.extra.data(), s.gtp.data(), s.on_tubule_cur.data())) ||
            (rc = mte_set_step(E.e, c.step)))
            return fail(E.e, "resume", rc);
        step = c.step;
    }
    Updater up(s, o);
    up.create_dcds();
    // DCD frames are written by a second host thread from a copy of the
    // output step's state while the engine runs the next segment; one frame
    // is in flight at a time, so the files keep the reference's frame order.
    FrameWriter writer(up);
    std::vector<mte_coord> r = s.coords, f(nb);
    std::vector<double> energies(7 * (size_t)nt);
    const long long stride = std::llabs(s.stride);
    const long long first = step;
    while (step < s.steps) {
        // A checkpoint holds the state at the top of its step, before that
        // step's list rebuild and events, so a resumed run replays them
        // exactly as the uninterrupted run does.
        if (o.checkpoint_every > 0 && step > 0 && step != first && step % o.checkpoint_every == 0) {
            Checkpoint c;
            c.step = step;
            c.rand_calls = g_rand_calls;
            c.coords.resize(nb);
            c.rng.resize(8 * nb);
            if ((rc = mte_get_state(E.e, c.coords.data())) || (rc = mte_get_rng(E.e, c.rng.data())))
                return fail(E.e, "checkpoint", rc);
            c.extra = s.extra;
            c.gtp = s.gtp;
            c.on_cur = s.on_tubule_cur;
            c.on_prev = s.on_tubule_prev;
            c.mt_len.assign(mt_len.begin(), mt_len.end());
            write_checkpoint(o.checkpoint_file, c, nt, n);
            if (!o.quiet) std::printf("Checkpoint at step %lld: %s\n", step, o.checkpoint_file.c_str());
        }
        // Lists first, as at the top of the reference's step
        // (src/compute_cuda.cu:1141-1152): the step's dynamics use lists of
        // the positions before hydrolysis and the concentration update,
        // whatever the output settings.
        if ((s.lj_on || s.is_assembly) && step % s.ljpairsupdatefreq == 0 && (rc = mte_rebuild_lists(E.e)))
            return fail(E.e, "mte_rebuild_lists", rc);
        {
            if (s.hydrolysis && s.hydrostep > 0 && step % s.hydrostep == 0 && step != 0) {
                up.hydrolyse();
                if ((rc = mte_set_flags(E.e, nullptr, s.gtp.data(), nullptr))) return fail(E.e, "mte_set_flags", rc);
            }
            if (step % stride == 0) {
                if (s.out_energy && (rc = mte_compute_energies(E.e, nullptr, energies.data())))
                    return fail(E.e, "mte_compute_energies", rc);
                if (s.out_force && (rc = mte_compute_forces(E.e, f.data()))) return fail(E.e, "mte_compute_forces", rc);
                if ((rc = mte_get_state(E.e, r.data()))) return fail(E.e, "mte_get_state", rc);
                auto update = [&]() {
                    if (!o.quiet) {
                        std::printf("Saving coordinates at step %lld\n", step);
                        print_time(step, s.steps);
                    }
                    writer.submit(r);
                    if (s.hydrolysis && step % (stride * 10) == 0) up.append_hydrolysis_pdb(step, r);
                    if (s.out_force) up.output_forces(f, r);
                    if (s.out_energy) up.output_energies(energies);
                };
                if (s.tub_length) {
                    s.on_tubule_prev = s.on_tubule_cur;
                    if (step != 0) {
                        mt_len_prev = mt_len;
                        if ((rc = up.mt_length(E.e, step, mt_len))) return fail(E.e, "mte_mt_length", rc);
                        if (s.is_const_conc) {
                            for (int t = 0; t < nt; t++) mt_len_prev[t] = mt_len[t] - mt_len_prev[t];
                            if (up.change_conc(mt_len_prev, mt_len, r)) {
                                if ((rc = mte_set_flags(E.e, s.extra.data(), nullptr, nullptr)))
                                    return fail(E.e, "mte_set_flags", rc);
                                if ((rc = mte_set_state(E.e, r.data()))) return fail(E.e, "mte_set_state", rc);
                            }
                        }
                        update();
                    } else {
                        update();
                        if ((rc = up.mt_length(E.e, step, mt_len))) return fail(E.e, "mte_mt_length", rc);
                    }
                } else {
                    update();
                }
            }
        }
        long long next = (step / stride + 1) * stride;
        if (s.hydrolysis && s.hydrostep > 0) next = std::min(next, (step / s.hydrostep + 1) * (long long)s.hydrostep);
        next = std::min(next, s.steps);
        if ((rc = mte_run(E.e, next - step))) return fail(E.e, "mte_run", rc);
        step = next;
    }
    writer.finish();
    // as the reference: the result PDBs hold the host copy of the last output step
    up.save_result_pdb(r);
    return MTE_OK;
}

}  // namespace mth
