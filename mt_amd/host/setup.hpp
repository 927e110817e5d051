// setup.hpp -- a run as the reference sets it up: the three configuration
// files read in the reference's order, the structure read from its PDB pair,
// the topology built from it, and the derived constants the engine takes.
//
//   load_run      initParameters, src/preparator.cpp:17-221 (keys and
//                 defaults of src/parameters.h:18-243), with read_PDB,
//                 :249-564, and AssemblyInit, :717-731
//   engine_params the mte_params subset of Parameters (mt_engine.h)
//   XYZ restarts  readRestart / writeRestart, :664-714, in the layout of
//                 src/xyzio.cpp
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "config.hpp"
#include "mt_engine.h"
#include "pdbio.hpp"

namespace mth {

constexpr float R_MON = 2.0f;   // src/mt.h:24
constexpr float R_MT = 8.12f;   // src/mt.h:23
constexpr float KB = 0.0019872041f;
constexpr int MAXLON = 8, MAXLAT = 16;   // src/preparator.cpp:721-722
constexpr int LAT_ZERO = 999999;         // src/mt.h:35

struct RunSetup {
    // main configuration file
    int device = 0;
    float dt = 0;
    int rseed = 0;
    long long steps = -1, stride = -1, firststep = 0;
    float ljpairscutoff = 0;
    int ljpairsupdatefreq = 0;
    std::string coord_xyz, coord_ang, ff_file, cond_file, restartkey;
    int fix = 1, ntr = 1;
    bool is_assembly = true, tub_length = true, out_energy = true, out_force = false, is_restart = false;
    bool tea_on = false, tea_capricious = true;
    float tea_a = R_MON, tea_epsmax = 999.f;
    int tea_epsilon_freq = 0;
    std::vector<std::string> dcd_xyz, dcd_ang, restart_xyz, restart_ang;
    // forcefield file
    float A_long = 0, A_lat = 0, D_long = 0, D_lat = 0, seam_coeff = 0;
    bool barrier = false;
    float a_barr_long = 0, r_barr_long = 0, w_barr_long = 0, a_barr_lat = 0, r_barr_lat = 0, w_barr_lat = 0;
    float C = 0, B_fi = 0, B_psi = 0, B_theta = 0, fi_0 = 0, psi_0 = 0, theta0_gdp = 0, theta0_gtp = 0;
    bool lj_on = false, is_wall = false;
    bool lj_regularize = false;   // the reference's compile-time REGULARIZATION, key "regularization"
    float ljscale = 0, ljsigma6 = 0, rep_leftborder = 0, rep_h = 0, rep_r = 0, rep_eps = 0;
    // conditions file
    bool is_const_conc = false, hydrolysis = false;
    float conc = 0, khydro = 0, Temp = 0, viscosity = 0, alpha = 0, freeze_temp = 0;
    long hydrostep = 0;
    // derived (preparator.cpp:205-219)
    float gammaR = 0, gammaTheta = 0, varR = 0, varTheta = 0;
    std::vector<float> zs;   // per trajectory wall height, starts at rep_h

    // structure: Ntot beads per trajectory, Ntr identical trajectories
    int ntot = 0;
    std::vector<PdbAtom> atoms;          // the xyz PDB records (names, chains, residues)
    std::vector<mte_coord> coords;       // Ntr * Ntot, trajectory-major
    // topology in the reference's encodings
    std::vector<int32_t> harmonic_count, harmonic;    // per monomer (shared by trajectories)
    std::vector<int32_t> mon_type;                    // per monomer
    std::vector<uint8_t> fixed;                       // per monomer
    std::vector<uint8_t> extra;                       // Ntr * Ntot
    std::vector<int32_t> gtp, on_tubule_cur, on_tubule_prev;   // Ntr * Ntot
    std::vector<int32_t> lon_count, lon, lat_count, lat;       // Ntr * Ntot (x MAXLON / MAXLAT)
};

// Reads the main configuration file `path` (with "name=value" overrides
// applied to every file, as src/configreader.cpp:60-74 does), then the
// forcefield and conditions files it names, and the PDB pair.  Relative
// file names are relative to the working directory, as in the reference.
RunSetup load_run(const std::string &path, const std::vector<std::string> &overrides);

// The engine's parameter block.
mte_params engine_params(const RunSetup &s);

// Angles wrapped toward 0 modulo 2 pi as at upload (src/compute_cuda.cu:1004-1010).
void wrap_angles(std::vector<mte_coord> &r);

// Rotated lateral attachment offsets p1, p2 of src/mt.h:44-50 (float as there).
void lateral_offsets(float p1[3], float p2[3]);

// XYZ files (src/xyzio.cpp): count line, comment line, "name x y z" rows.
struct XyzAtom { char name; double x, y, z; };
std::vector<XyzAtom> read_xyz(const std::string &path);
void write_xyz(const std::string &path, const std::vector<XyzAtom> &atoms);
// readRestart / writeRestart: per trajectory an xyz file of positions and
// one of angles (fi, psi, theta), plus the key file holding the step.
void read_restart(RunSetup &s);
void write_restart(const RunSetup &s, const std::vector<mte_coord> &r, long long step);

}  // namespace mth
