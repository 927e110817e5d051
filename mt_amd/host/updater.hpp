// updater.hpp -- the reference's host-side updater and its step loop, driving
// the engine through the C ABI.
//
//   Updater::output_energies   OutputAllEnergies, src/updater.cpp:3-43
//   Updater::mt_length         mt_length, :154-217 (on_tubule_cur, mt_len.dat; the
//                              classification itself runs on the device, mte_mt_length)
//   Updater::change_conc       change_conc, :100-152 (glibc rand(), as there)
//   Updater::hydrolyse         hydrolyse, :219-256 (glibc rand())
//   Updater::update            update, :74-98 (DCD frames, energies, forces)
//   run                        compute(), src/compute_cuda.cu:1125-1260, and
//                              main(), src/main.cpp:55-88: events at every
//                              stride and hydrolysis step, mte_run between them
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mt_engine.h"
#include "pdbio.hpp"
#include "setup.hpp"

namespace mth {

struct RunOptions {
    int32_t mode = MTE_MODE_AUTO;   // engine execution mode
    int32_t max_lj = 256;           // LJ list capacity (the reference's)
    bool quiet = false;             // suppress the per-event log lines
    int autotune_steps = 0;         // > 0: mte_autotune with this many steps per trial before the run
    long long checkpoint_every = 0; // > 0: write a checkpoint at every multiple of this many steps
    std::string checkpoint_file = "checkpoint.bin";
    std::string resume_file;        // non-empty: continue the run from this checkpoint
    std::string result_xyz = "result_xyz.pdb", result_ang = "result_ang.pdb";
    std::string mt_len_file = "mt_len.dat", hydrolysis_pdb = "dcd/hydrolysis.pdb";
};

class Updater {
public:
    Updater(RunSetup &s, const RunOptions &o);
    void output_energies(const std::vector<double> &per_traj) const;
    void output_forces(const std::vector<mte_coord> &f, const std::vector<mte_coord> &r) const;
    int mt_length(mte_engine *e, long long step, std::vector<int> &mt_len);
    int change_conc(const std::vector<int> &delta, const std::vector<int> &mt_len, std::vector<mte_coord> &r);
    void hydrolyse();
    void create_dcds() const;
    void save_dcd(const std::vector<mte_coord> &r) const;
    void append_hydrolysis_pdb(long long step, const std::vector<mte_coord> &r) const;
    void save_result_pdb(const std::vector<mte_coord> &r) const;

private:
    RunSetup &s_;
    RunOptions o_;
    DcdWriter dcd_;
};

// The whole run: srand(rseed), angle wrap, engine creation, the step loop
// with its host events, the result PDBs.  Returns 0 or an MTE_ERR_* code
// (messages on stderr).
int run(RunSetup &s, const RunOptions &o);

}  // namespace mth
