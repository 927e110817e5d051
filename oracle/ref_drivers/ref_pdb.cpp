// Driver (this repo's own code) around the reference's PDB reader,
// src/pdbio.cpp, compiled where it lies.  Prints one line per ATOM record:
// id name chain resid x y z occupancy beta.
#include <cstdio>
#include "pdbio.h"

int main(int argc, char** argv) {
    if (argc != 2) { std::fprintf(stderr, "usage: ref_pdb FILE\n"); return 2; }
    PDB pdb;
    readPDB(argv[1], &pdb);
    for (int i = 0; i < pdb.atomCount; i++) {
        const PDBAtom& a = pdb.atoms[i];
        std::fprintf(stderr, "@@%d %s %c %d %.6f %.6f %.6f %.6f %.6f\n", a.id, a.name, a.chain, a.resid,
                     a.x, a.y, a.z, a.occupancy, a.beta);
    }
    return 0;
}
