// Driver (this repo's own code) around the reference's XYZ reader and
// writer, src/xyzio.cpp (+ src/wrapper.cpp), compiled where they lie.
// usage: ref_xyz IN OUT -- reads IN with readXYZ, prints "name x y z" rows
// ("%.9g") on stderr and writes OUT with writeXYZ.
#include <cstdio>
#include "xyzio.h"

int main(int argc, char** argv) {
    if (argc != 3) { std::fprintf(stderr, "usage: ref_xyz IN OUT\n"); return 2; }
    XYZ xyz;
    readXYZ(argv[1], &xyz);
    for (int i = 0; i < xyz.atomCount; i++)
        std::fprintf(stderr, "@@%c %.9g %.9g %.9g\n", xyz.atoms[i].name, xyz.atoms[i].x, xyz.atoms[i].y,
                     xyz.atoms[i].z);
    writeXYZ(argv[2], &xyz);
    return 0;
}
