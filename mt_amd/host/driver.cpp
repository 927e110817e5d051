// driver.cpp -- `mt_host CONFIG [name=value ...]`: the reference's command
// line (src/main.cpp:25-88) on the engine.  Reads the three configuration
// files and the PDB pair, runs the steps with the reference's output cadence
// (DCD frames, energies, mt_len.dat, hydrolysis and constant-concentration
// updates), and writes result_xyz.pdb / result_ang.pdb.
//
// Keys read in addition to the reference's (main file, all optional):
//   engine_mode    auto | streamed | resident   (default auto)
//   engine_max_lj  LJ list capacity per bead     (default 256, the reference's)
//   engine_quiet   yes | no                      (default no)
//   engine_autotune  steps per trial of mte_autotune before the run (default 0: off)
#include <cstdio>
#include <string>
#include <vector>

#include "config.hpp"
#include "setup.hpp"
#include "updater.hpp"

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s CONFIG [name=value ...]\n", argv[0]);
        return 2;
    }
    std::vector<std::string> overrides(argv + 2, argv + argc);
    try {
        mth::Config main_cfg;
        main_cfg.parse_file(argv[1], overrides);
        mth::RunOptions opt;
        const std::string mode = main_cfg.masked("engine_mode", "auto");
        opt.mode = mode == "resident" ? MTE_MODE_RESIDENT : mode == "streamed" ? MTE_MODE_STREAMED : MTE_MODE_AUTO;
        opt.max_lj = main_cfg.integer("engine_max_lj", 256);
        opt.quiet = main_cfg.yesno("engine_quiet", false);
        opt.autotune_steps = main_cfg.integer("engine_autotune", 0);
        mth::RunSetup setup = mth::load_run(argv[1], overrides);
        return mth::run(setup, opt) == MTE_OK ? 0 : 1;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
}
