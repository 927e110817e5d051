// Driver (this repo's own code) around the reference's DCD writer,
// src/dcdio.cpp, compiled where it lies.
// usage: ref_dcd OUT NATOMS NFRAMES_HEADER FIRST DT FREQ < frames
// stdin: whitespace-separated floats, per frame X[0..N) Y[0..N) Z[0..N);
// writes the header, then every frame read, as compute() does through
// createDCD / dcdOpenWrite / dcdWriteHeader / dcdOpenAppend / dcdWriteFrame.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "dcdio.h"

int main(int argc, char** argv) {
    if (argc != 7) { std::fprintf(stderr, "usage: ref_dcd OUT N NFRAMES FIRST DT FREQ\n"); return 2; }
    const int n = std::atoi(argv[2]);
    DCD dcd;
    createDCD(&dcd, n, std::atoi(argv[3]), std::atoi(argv[4]), (float)std::atof(argv[5]), std::atoi(argv[6]), 0, 0,
              0, 0);
    dcdOpenWrite(&dcd, argv[1]);
    dcdWriteHeader(dcd);
    dcdClose(dcd);
    std::vector<float> buf(3 * (size_t)n);
    for (;;) {
        size_t k = 0;
        for (; k < buf.size(); k++)
            if (std::scanf("%f", &buf[k]) != 1) break;
        if (k < buf.size()) break;
        for (int i = 0; i < n; i++) {
            dcd.frame.X[i] = buf[i];
            dcd.frame.Y[i] = buf[n + i];
            dcd.frame.Z[i] = buf[2 * n + i];
        }
        dcdOpenAppend(&dcd, argv[1]);
        dcdWriteFrame(dcd);
        dcdClose(dcd);
    }
    return 0;
}
