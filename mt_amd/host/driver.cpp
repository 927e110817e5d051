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
//   engine_checkpoint  write a checkpoint every this many steps (default 0: off;
//                  a multiple of stride, LJPairsUpdateFreq, tea_epsilon_freq and the
//                  hydrolysis step, so that nothing is pending at a checkpoint)
//   engine_checkpoint_file  its path (default checkpoint.bin)
//   engine_resume  continue from this checkpoint: state, HybridTaus streams, step,
//                  flags and the host rand() sequence as they were, so the steps
//                  after it are those of the uninterrupted run
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
        opt.checkpoint_every = main_cfg.long_integer("engine_checkpoint", 0);
        opt.checkpoint_file = main_cfg.masked("engine_checkpoint_file", "checkpoint.bin");
        if (main_cfg.has("engine_resume")) opt.resume_file = main_cfg.masked("engine_resume");
        mth::RunSetup setup = mth::load_run(argv[1], overrides);
        return mth::run(setup, opt) == MTE_OK ? 0 : 1;
    } catch (const std::exception &ex) {
        std::fprintf(stderr, "%s\n", ex.what());
        return 1;
    }
}
