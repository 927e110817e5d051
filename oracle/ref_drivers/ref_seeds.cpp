// Driver (this repo's own code) around the reference's ran2() from
// src/ran2.h, compiled where it lies.  Prints the 4*np seed words that
// generateSeeds (src/HybridTaus.cu:32-48) derives from `seed`, one per line.
// ran2 keeps its state in function statics, so each run starts fresh.
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include "ran2.h"

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_seeds SEED NP\n"); return 2; }
    int seed = std::atoi(argv[1]);
    long np = std::atol(argv[2]);
    for (long i = 0; i < 4 * np; i++) {
        unsigned v;
        do { v = (unsigned)(ran2(&seed) * UINT_MAX); } while (v < 128);
        std::printf("%u\n", v);
    }
    return 0;
}
