// capi.cpp -- C entry points of the host library (libmt_host.so) for
// bindings and tests: configuration queries, PDB parsing, run setup and the
// whole run.  Strings in, plain arrays out; errors as codes with a message
// from mth_last_error().
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "config.hpp"
#include "pdbio.hpp"
#include "setup.hpp"
#include "updater.hpp"

namespace {
thread_local std::string g_err;

int caught(const std::exception &ex) {
    g_err = ex.what();
    return 1;
}

std::vector<std::string> to_vec(const char *const *v, int n) {
    std::vector<std::string> out;
    for (int i = 0; i < n; i++) out.emplace_back(v[i]);
    return out;
}

int put(const std::string &s, char *out, int outlen) {
    if ((int)s.size() + 1 > outlen) {
        g_err = "output buffer too small";
        return 2;
    }
    std::memcpy(out, s.c_str(), s.size() + 1);
    return 0;
}
}  // namespace

extern "C" {

const char *mth_last_error(void) { return g_err.c_str(); }

// One query against a parsed file: kind is str | masked | int | long | float
// | yesno; def NULL means required.  The value comes back as text (numbers
// printed as "%d", "%lld", "%.9g", yes/no as 1/0).
int mth_config_query(const char *path, const char *const *overrides, int n_over, const char *kind,
                     const char *name, const char *def, char *out, int outlen) {
    try {
        mth::Config c;
        c.parse_file(path, to_vec(overrides, n_over));
        const std::string k = kind, nm = name;
        char buf[128];
        std::string v;
        if (k == "str") {
            v = def ? c.str(nm, def) : c.str(nm);
        } else if (k == "masked") {
            v = def ? c.masked(nm, def) : c.masked(nm);
        } else if (k == "int") {
            std::snprintf(buf, sizeof buf, "%d", def ? c.integer(nm, std::atoi(def)) : c.integer(nm));
            v = buf;
        } else if (k == "long") {
            std::snprintf(buf, sizeof buf, "%lld", def ? c.long_integer(nm, std::atol(def)) : c.long_integer(nm));
            v = buf;
        } else if (k == "float") {
            std::snprintf(buf, sizeof buf, "%.9g", (double)(def ? c.real(nm, (float)std::atof(def)) : c.real(nm)));
            v = buf;
        } else if (k == "yesno") {
            v = (def ? c.yesno(nm, std::atoi(def) != 0) : c.yesno(nm)) ? "1" : "0";
        } else {
            g_err = "unknown query kind";
            return 2;
        }
        return put(v, out, outlen);
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

// ATOM records as text, one "id name chain resid x y z occupancy beta" line
// each ("%.6f" numbers).
int mth_pdb_dump(const char *path, char *out, int outlen) {
    try {
        std::string s;
        char line[256];
        for (const mth::PdbAtom &a : mth::read_pdb(path)) {
            std::snprintf(line, sizeof line, "%d %s %c %d %.6f %.6f %.6f %.6f %.6f\n", a.id, a.name.c_str(), a.chain,
                          a.resid, a.x, a.y, a.z, a.occupancy, a.beta);
            s += line;
        }
        return put(s, out, outlen);
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

// Round trip: read `in`, write it with the PDB writer to `out_path`.
int mth_pdb_rewrite(const char *in, const char *out_path) {
    try {
        mth::write_pdb(out_path, mth::read_pdb(in));
        return 0;
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

// XYZ round trip: read `in` (rows as "name x y z" with "%.9g" into `dump`),
// write it to `out_path` with the XYZ writer.
int mth_xyz_rewrite(const char *in, const char *out_path, char *dump, int dumplen) {
    try {
        const std::vector<mth::XyzAtom> a = mth::read_xyz(in);
        std::string s;
        char line[256];
        for (const mth::XyzAtom &x : a) {
            std::snprintf(line, sizeof line, "%c %.9g %.9g %.9g\n", x.name, x.x, x.y, x.z);
            s += line;
        }
        mth::write_xyz(out_path, a);
        return put(s, dump, dumplen);
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

int mth_dcd_write(const char *path, int natoms, int nframes_header, int first, float dt, int freq, int nframes,
                  const float *xyz) {
    try {
        mth::DcdWriter w(natoms, nframes_header, first, dt, freq);
        w.create(path);
        for (int f = 0; f < nframes; f++) {
            const float *p = xyz + (size_t)f * 3 * natoms;
            w.append(path, p, p + natoms, p + 2 * natoms);
        }
        return 0;
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

struct mth_setup {
    mth::RunSetup s;
};

int mth_setup_load(const char *path, const char *const *overrides, int n_over, mth_setup **out) {
    try {
        auto *h = new mth_setup{mth::load_run(path, to_vec(overrides, n_over))};
        *out = h;
        return 0;
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

void mth_setup_free(mth_setup *h) { delete h; }

int mth_setup_sizes(const mth_setup *h, int32_t *ntot, int32_t *ntr, int64_t *steps, int64_t *stride) {
    *ntot = h->s.ntot;
    *ntr = h->s.ntr;
    *steps = h->s.steps;
    *stride = h->s.stride;
    return 0;
}

int mth_setup_params(const mth_setup *h, mte_params *p) {
    *p = mth::engine_params(h->s);
    return 0;
}

// Per-monomer harmonic count / partner / mon_type / fixed (Ntot), per-bead
// extra, lists and coordinates (Ntr * Ntot), as read (angles not yet wrapped).
int mth_setup_topology(const mth_setup *h, int32_t *harmonic_count, int32_t *harmonic, int32_t *mon_type,
                       uint8_t *fixed, uint8_t *extra, int32_t *lon_count, int32_t *lon, int32_t *lat_count,
                       int32_t *lat, mte_coord *coords) {
    const mth::RunSetup &s = h->s;
    std::copy(s.harmonic_count.begin(), s.harmonic_count.end(), harmonic_count);
    std::copy(s.harmonic.begin(), s.harmonic.end(), harmonic);
    std::copy(s.mon_type.begin(), s.mon_type.end(), mon_type);
    std::copy(s.fixed.begin(), s.fixed.end(), fixed);
    std::copy(s.extra.begin(), s.extra.end(), extra);
    std::copy(s.lon_count.begin(), s.lon_count.end(), lon_count);
    std::copy(s.lon.begin(), s.lon.end(), lon);
    std::copy(s.lat_count.begin(), s.lat_count.end(), lat_count);
    std::copy(s.lat.begin(), s.lat.end(), lat);
    std::copy(s.coords.begin(), s.coords.end(), coords);
    return 0;
}

// The whole run of `path` (what mt_host does), quiet or not.
int mth_run(const char *path, const char *const *overrides, int n_over, int32_t mode, int32_t max_lj, int32_t quiet) {
    try {
        mth::RunSetup s = mth::load_run(path, to_vec(overrides, n_over));
        mth::RunOptions o;
        o.mode = mode;
        o.max_lj = max_lj;
        o.quiet = quiet != 0;
        mth::Config main_cfg;   // the checkpoint keys, as mt_host reads them
        main_cfg.parse_file(path, to_vec(overrides, n_over));
        o.checkpoint_every = main_cfg.long_integer("engine_checkpoint", 0);
        o.checkpoint_file = main_cfg.masked("engine_checkpoint_file", "checkpoint.bin");
        if (main_cfg.has("engine_resume")) o.resume_file = main_cfg.masked("engine_resume");
        const int rc = mth::run(s, o);
        if (rc) g_err = "run failed with engine status " + std::to_string(rc);
        return rc;
    } catch (const std::exception &ex) {
        return caught(ex);
    }
}

}  // extern "C"
