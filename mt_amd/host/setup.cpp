// setup.cpp -- see setup.hpp.
#include "setup.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mth {

namespace {

// Rotation R = Rz(psi) Ry(theta) Rx(fi), the matrix whose entries the
// reference expands (src/compute_cuda.cu:89-96, src/preparator.cpp:536-548).
void rotate(double fi, double theta, double psi, const float v[3], double out[3]) {
    const double cf = std::cos(fi), sf = std::sin(fi), ct = std::cos(theta), st = std::sin(theta),
                 cp = std::cos(psi), sp = std::sin(psi);
    const double R[3][3] = {{cp * ct, cp * st * sf - sp * cf, cp * st * cf + sp * sf},
                            {sp * ct, sp * st * sf + cp * cf, sp * st * cf - cp * sf},
                            {-st, ct * sf, ct * cf}};
    for (int r = 0; r < 3; r++) out[r] = R[r][0] * v[0] + R[r][1] * v[1] + R[r][2] * v[2];
}

// read_PDB, src/preparator.cpp:249-356: per-bead records, harmonic partners.
void read_structure(RunSetup &s) {
    s.atoms = read_pdb(s.coord_xyz);
    const std::vector<PdbAtom> ang = read_pdb(s.coord_ang);
    const int n = (int)s.atoms.size();
    if (n == 0) throw IoError(s.coord_xyz + ": no ATOM records");
    if ((int)ang.size() < n) throw IoError(s.coord_ang + ": fewer ATOM records than " + s.coord_xyz);
    s.ntot = n;
    const size_t nb = (size_t)n * s.ntr;
    s.mon_type.assign(n, 0);
    s.fixed.assign(n, 0);
    for (int i = 0; i < n; i++) {
        const std::string &nm = s.atoms[i].name;
        if (nm.size() > 1 && nm[1] == 'B') s.mon_type[i] = 1;       // 'A' -> 0 (:259-265)
        s.fixed[i] = s.atoms[i].resid <= s.fix ? 1 : 0;             // :280-286
    }
    s.extra.assign(nb, 0);
    s.gtp.assign(nb, 0);
    s.on_tubule_cur.assign(nb, 0);
    s.on_tubule_prev.assign(nb, 0);
    s.coords.assign(nb, mte_coord{});
    for (int t = 0; t < s.ntr; t++)
        for (int i = 0; i < n; i++) {
            const size_t p = (size_t)t * n + i;
            s.extra[p] = s.atoms[i].chain == 'X' ? 1 : 0;            // :288-301
            mte_coord &c = s.coords[p];
            c.x = (float)s.atoms[i].x;
            c.y = (float)s.atoms[i].y;
            c.z = (float)s.atoms[i].z;
            c.fi = (float)ang[i].x;                                  // angle PDB: fi, psi, theta in x, y, z
            c.psi = (float)ang[i].y;
            c.theta = (float)ang[i].z;
            c.w = 0.f;
        }
    // covalent (harmonic) partners: same resid and chain, |i - j| < 2 (:325-355)
    s.harmonic_count.assign(n, 0);
    s.harmonic.assign(n, -1);
    for (int i = 0; i < n; i++)
        for (int j = std::max(0, i - 1); j <= std::min(n - 1, i + 1); j++) {
            if (j == i || s.atoms[i].resid != s.atoms[j].resid || s.atoms[i].chain != s.atoms[j].chain) continue;
            if (s.harmonic_count[i] >= 1)
                throw IoError("bead " + std::to_string(i) + " has two harmonic partners; the engine supports one");
            s.harmonic[i] = s.atoms[i].id > s.atoms[j].id ? j : -j;
            s.harmonic_count[i]++;
        }
    s.lon_count.assign(nb, 0);
    s.lon.assign(nb * MAXLON, 0);
    s.lat_count.assign(nb, 0);
    s.lat.assign(nb * MAXLAT, 0);
    if (s.is_assembly) return;   // AssemblyInit (:717-731): empty lists, rebuilt on the device
    // static lists from trajectory 0's coordinates (:357-561), shared by every trajectory
    float p1[3], p2[3];
    lateral_offsets(p1, p2);
    std::vector<int32_t> lc(n, 0), l(n * (size_t)MAXLON, 0), ac(n, 0), a(n * (size_t)MAXLAT, 0);
    for (int i = 0; i < n; i++) {
        if (s.extra[i]) continue;
        for (int j = 0; j < n; j++) {
            if (std::abs(s.atoms[i].resid - s.atoms[j].resid) == 1 && s.atoms[i].chain == s.atoms[j].chain &&
                std::abs(s.atoms[i].id - s.atoms[j].id) == 1) {
                if (lc[i] >= MAXLON) throw IoError("longitudinal list overflow at bead " + std::to_string(i));
                l[(size_t)i * MAXLON + lc[i]++] = s.atoms[i].id > s.atoms[j].id ? j : -j;
            }
        }
    }
    for (int i = 0; i < n; i++) {
        if (s.extra[i]) continue;
        const mte_coord &ci = s.coords[i];
        for (int j = 0; j < n; j++) {
            const mte_coord &cj = s.coords[j];
            for (int ind = 0; ind < 2; ind++) {
                // ind 0: bead i carries p1, bead j p2; ind 1 the other way round
                double ai[3], bj[3];
                rotate(ci.fi, ci.theta, ci.psi, ind == 0 ? p1 : p2, ai);
                rotate(cj.fi, cj.theta, cj.psi, ind == 0 ? p2 : p1, bj);
                const double gx = (double)ci.x - cj.x + ai[0] - bj[0], gy = (double)ci.y - cj.y + ai[1] - bj[1],
                             gz = (double)ci.z - cj.z + ai[2] - bj[2];
                if (std::sqrt(gx * gx + gy * gy + gz * gz) < R_MON) {
                    if (ac[i] >= MAXLAT) throw IoError("lateral list overflow at bead " + std::to_string(i));
                    a[(size_t)i * MAXLAT + ac[i]++] = ind == 1 ? j : -j;
                }
            }
        }
    }
    for (int t = 0; t < s.ntr; t++) {
        std::copy(lc.begin(), lc.end(), s.lon_count.begin() + (size_t)t * n);
        std::copy(l.begin(), l.end(), s.lon.begin() + (size_t)t * n * MAXLON);
        std::copy(ac.begin(), ac.end(), s.lat_count.begin() + (size_t)t * n);
        std::copy(a.begin(), a.end(), s.lat.begin() + (size_t)t * n * MAXLAT);
    }
}

}  // namespace

void lateral_offsets(float p1[3], float p2[3]) {
    const float ang = (float)(2.0 * M_PI / 13.0);
    const float x = 0.5f * R_MT * (std::cos(ang) - 1.0f), y = 0.5f * R_MT * std::sin(ang);
    p1[0] = x; p1[1] = y; p1[2] = -3.0f * R_MON / 13.0f;
    p2[0] = x; p2[1] = -y; p2[2] = 3.0f * R_MON / 13.0f;
}

RunSetup load_run(const std::string &path, const std::vector<std::string> &ov) {
    RunSetup s;
    Config c;
    // main file (src/preparator.cpp:20-113)
    c.parse_file(path, ov);
    s.device = c.integer("device");
    s.dt = c.real("dt");
    s.rseed = c.integer("rseed");
    s.steps = c.long_integer("steps", -1);
    s.stride = c.long_integer("stride", -1);
    s.ljpairscutoff = c.real("LJPairsCutoff");
    s.ljpairsupdatefreq = c.integer("LJPairsUpdateFreq");
    s.coord_ang = c.masked("coordinates_ang");
    s.coord_xyz = c.masked("coordinates_xyz");
    s.ff_file = c.masked("forcefield");
    s.cond_file = c.masked("conditions");
    s.fix = c.integer("fix", 1);
    s.ntr = c.integer("runnum", 1);
    if (s.ntr < 1) throw ConfigError("runnum must be at least 1");
    s.is_assembly = c.yesno("is_assembly", true);
    s.tub_length = c.yesno("tubule_length", true);
    s.out_energy = c.yesno("output_energy", true);
    s.out_force = c.yesno("output_force", false);
    s.tea_on = c.yesno("tea_on", false);
    if (s.tea_on) {
        s.tea_a = c.real("tea_a", R_MON);
        s.tea_capricious = c.yesno("tea_capricious", true);
        s.tea_epsilon_freq = c.integer("tea_epsilon_freq");
        s.tea_epsmax = c.real("tea_epsmax", 999.f);
    }
    s.restartkey = c.masked("restartkey");
    s.is_restart = c.yesno("is_restart", false);
    if (!s.is_restart) {
        s.firststep = c.long_integer("firststep", 0);
    } else {
        std::ifstream key(s.restartkey);
        if (!(key >> s.firststep)) throw IoError("Reading restartkey " + s.restartkey + ": unable to get firststep");
    }
    read_structure(s);
    for (int t = 0; t < s.ntr; t++) {
        const std::string run = std::to_string(t);
        s.dcd_ang.push_back(c.masked_replace("dcd_ang", run, "<run>"));
        s.dcd_xyz.push_back(c.masked_replace("dcd_xyz", run, "<run>"));
        s.restart_xyz.push_back(c.masked_replace("restart_xyz", run, "<run>"));
        s.restart_ang.push_back(c.masked_replace("restart_ang", run, "<run>"));
    }
    if (s.is_restart) read_restart(s);

    // forcefield file (:121-178)
    c.parse_file(s.ff_file, ov);
    s.A_long = c.real("A_long");
    s.A_lat = c.real("A_lat");
    s.D_long = c.real("D_long");
    s.D_lat = c.real("D_lat");
    s.seam_coeff = c.real("seam_coeff");
    const double fwhm = 2 * std::sqrt(-2 * std::log(0.5));
    s.barrier = c.yesno("barrier");
    if (s.barrier) {
        s.a_barr_long = c.real("a_barr_long");
        s.r_barr_long = c.real("r_barr_long");
        s.w_barr_long = (float)(c.real("w_barr_long") / fwhm);
        s.a_barr_lat = c.real("a_barr_lat");
        s.r_barr_lat = c.real("r_barr_lat");
        s.w_barr_lat = (float)(c.real("w_barr_lat") / fwhm);
    }
    s.C = c.real("C");
    s.B_fi = c.real("B_fi");
    s.B_psi = c.real("B_psi");
    s.B_theta = c.real("B_theta");
    s.fi_0 = c.real("fi0");
    s.psi_0 = c.real("psi0");
    s.theta0_gdp = c.real("theta0_gdp");
    s.theta0_gtp = c.real("theta0_gtp");
    s.lj_on = c.yesno("LJ_on");
    if (s.lj_on) {
        s.ljscale = c.real("LJScale");
        s.ljsigma6 = (float)std::pow(c.real("LJSigma"), 6);
    }
    s.lj_regularize = c.yesno("regularization", false);
    s.is_wall = c.yesno("repulsive_walls");
    s.rep_leftborder = c.real("rep_leftborder");
    s.rep_h = c.real("rep_h");
    s.rep_r = c.real("rep_r");
    s.rep_eps = c.real("rep_eps");
    s.zs.assign(s.ntr, s.rep_h);

    // conditions file (:180-221)
    c.parse_file(s.cond_file, ov);
    s.is_const_conc = c.yesno("is_const_conc");
    if (s.is_const_conc) s.conc = c.real("conc");
    s.hydrolysis = c.yesno("hydrolysis");
    if (s.hydrolysis) {
        s.khydro = c.real("khydro");
        s.hydrostep = (long)(0.02 * 1000000000000LL / (double)(s.dt * s.khydro));
    }
    std::fill(s.gtp.begin(), s.gtp.end(), 1);
    s.Temp = c.real("Temp");
    s.viscosity = c.real("viscosity");
    s.gammaR = (float)(6 * M_PI * s.viscosity * (s.tea_on ? s.tea_a : R_MON));
    s.gammaTheta = (float)(8 * M_PI * s.viscosity * std::pow(R_MON, 3));
    s.varR = std::sqrt(2.0f * KB * s.Temp * s.dt / s.gammaR);
    s.varTheta = std::sqrt(2.0f * KB * s.Temp * s.dt / s.gammaTheta);
    s.alpha = c.real("alpha");
    s.freeze_temp = c.real("freeze_temp");
    int extra_count = 0;
    for (int i = 0; i < s.ntot; i++) extra_count += s.extra[i];
    if (extra_count == 0 && s.is_const_conc)
        throw ConfigError("Error! You want constant concentration! Load structure with extra particles (chain X) "
                          "to support constant concentration!");
    return s;
}

mte_params engine_params(const RunSetup &s) {
    mte_params p{};
    p.dt = s.dt;
    p.temp = s.Temp;
    p.gammaR = s.gammaR;
    p.gammaTheta = s.gammaTheta;
    p.varR = s.varR;
    p.varTheta = s.varTheta;
    p.alpha = s.alpha;
    p.freeze_temp = s.freeze_temp;
    p.C = s.C;
    p.B_psi = s.B_psi;
    p.B_fi = s.B_fi;
    p.B_theta = s.B_theta;
    p.psi_0 = s.psi_0;
    p.fi_0 = s.fi_0;
    p.theta0_gtp = s.theta0_gtp;
    p.theta0_gdp = s.theta0_gdp;
    p.A_long = s.A_long;
    p.D_long = s.D_long;
    p.A_lat = s.A_lat;
    p.D_lat = s.D_lat;
    p.seam_coeff = s.seam_coeff;
    p.barrier = s.barrier;
    p.a_barr_long = s.a_barr_long;
    p.r_barr_long = s.r_barr_long;
    p.w_barr_long = s.w_barr_long;
    p.a_barr_lat = s.a_barr_lat;
    p.r_barr_lat = s.r_barr_lat;
    p.w_barr_lat = s.w_barr_lat;
    p.lj_on = s.lj_on;
    p.ljscale = s.ljscale;
    p.ljsigma6 = s.ljsigma6;
    p.ljpairscutoff = s.ljpairscutoff;
    p.ljpairsupdatefreq = s.ljpairsupdatefreq;
    p.is_wall = s.is_wall;
    p.rep_h = s.rep_h;
    p.rep_r = s.rep_r;
    p.rep_eps = s.rep_eps;
    p.rep_leftborder = s.rep_leftborder;
    p.is_assembly = s.is_assembly;
    p.tea_on = s.tea_on;
    p.tea_a = s.tea_a;
    p.tea_epsilon_freq = s.tea_on ? s.tea_epsilon_freq : 1;
    p.tea_capricious = s.tea_capricious;
    p.tea_epsmax = s.tea_epsmax;
    p.lj_regularize = s.lj_regularize;
    return p;
}

void wrap_angles(std::vector<mte_coord> &r) {
    const double two_pi = 2 * M_PI;
    for (mte_coord &c : r) {
        c.fi = (float)(c.fi - two_pi * (int)(c.fi / two_pi));
        c.psi = (float)(c.psi - two_pi * (int)(c.psi / two_pi));
        c.theta = (float)(c.theta - two_pi * (int)(c.theta / two_pi));
    }
}

std::vector<XyzAtom> read_xyz(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw IoError("Error opening '" + path + "'");
    std::string line;
    std::getline(in, line);
    const int n = std::atoi(line.c_str());
    std::getline(in, line);
    std::vector<XyzAtom> out(n > 0 ? n : 0);
    for (int i = 0; i < n; i++) {
        if (!std::getline(in, line)) throw IoError(path + ": truncated");
        std::istringstream ss(line);
        std::string name, x, y, z;
        if (!(ss >> name >> x >> y >> z)) throw IoError(path + ": bad row " + std::to_string(i));
        out[i] = XyzAtom{name[0], std::atof(x.c_str()), std::atof(y.c_str()), std::atof(z.c_str())};
    }
    return out;
}

void write_xyz(const std::string &path, const std::vector<XyzAtom> &atoms) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw IoError("Error opening '" + path + "'");
    std::fprintf(f, "%d\n", (int)atoms.size());
    std::fprintf(f, "Created by 'xyzio.cpp'\n");
    for (const XyzAtom &a : atoms) std::fprintf(f, "%-*c%*f%*f%*f\n", 16, a.name, 16, a.x, 16, a.y, 16, a.z);
    std::fclose(f);
}

void read_restart(RunSetup &s) {
    for (int t = 0; t < s.ntr; t++) {
        const std::vector<XyzAtom> xyz = read_xyz(s.restart_xyz[t]);
        const std::vector<XyzAtom> ang = read_xyz(s.restart_ang[t]);
        if ((int)xyz.size() < s.ntot || (int)ang.size() < s.ntot)
            throw IoError("restart files of trajectory " + std::to_string(t) + " hold fewer atoms than the PDB");
        for (int i = 0; i < s.ntot; i++) {
            mte_coord &c = s.coords[(size_t)t * s.ntot + i];
            c.x = (float)xyz[i].x;
            c.y = (float)xyz[i].y;
            c.z = (float)xyz[i].z;
            c.fi = (float)ang[i].x;
            c.psi = (float)ang[i].y;
            c.theta = (float)ang[i].z;
        }
    }
}

void write_restart(const RunSetup &s, const std::vector<mte_coord> &r, long long step) {
    std::vector<XyzAtom> a(s.ntot);
    for (int t = 0; t < s.ntr; t++) {
        for (int i = 0; i < s.ntot; i++) {
            const mte_coord &c = r[(size_t)t * s.ntot + i];
            a[i] = XyzAtom{s.atoms[i].name[0], c.x, c.y, c.z};
        }
        write_xyz(s.restart_xyz[t], a);
        for (int i = 0; i < s.ntot; i++) {
            const mte_coord &c = r[(size_t)t * s.ntot + i];
            a[i] = XyzAtom{s.atoms[i].name[0], c.fi, c.psi, c.theta};
        }
        write_xyz(s.restart_ang[t], a);
    }
    FILE *k = std::fopen(s.restartkey.c_str(), "w");
    if (!k) throw IoError("cannot write restart key " + s.restartkey);
    std::fprintf(k, "%lld", step);
    std::fclose(k);
}

}  // namespace mth
