// pdbio.hpp -- PDB ATOM records and DCD trajectories in the layouts the
// reference reads and writes (src/pdbio.cpp, src/dcdio.cpp), so that its
// input files and analysis tools work unchanged.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace mth {

struct IoError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct PdbAtom {
    int id = 0;
    std::string name;      // first word of columns 13-16
    char altLoc = ' ';
    std::string resName;   // columns 18-20, untrimmed
    char chain = ' ';
    int resid = 0;
    double x = 0, y = 0, z = 0, occupancy = 0, beta = 0;
};

// ATOM records of a PDB file (src/pdbio.cpp:36-79, fields :181-214): lines
// are read in fgets chunks of 79 characters; a record is a chunk whose first
// space-delimited token is "ATOM"; numeric fields by column with atoi/atof.
std::vector<PdbAtom> read_pdb(const std::string &path);

// writePDB (src/pdbio.cpp:276-320): serials i + 1 (hexadecimal from 100000
// atoms on), "%8.3f" coordinates, "%6.2f" occupancy and beta, then "END".
void write_pdb(const std::string &path, const std::vector<PdbAtom> &atoms);

// CHARMM/NAMD DCD writer with the reference's header (src/dcdio.cpp:16-131):
// NFILE frames announced, NPRIV first frame, NSAVC stride, NPRIV-NSAVC in the
// NSTEP slot, DELTA, no unit cell, two 80-byte remarks.
class DcdWriter {
public:
    DcdWriter(int natoms, int nframes, int first_frame, float timestep, int save_freq);
    // Truncates `path` and writes the header (dcdOpenWrite + dcdWriteHeader).
    void create(const std::string &path) const;
    // Appends one frame (dcdOpenAppend + dcdWriteFrame).
    void append(const std::string &path, const float *x, const float *y, const float *z) const;

private:
    int n_, nfile_, npriv_, nsavc_;
    float delta_;
    std::string remark1_, remark2_;
};

}  // namespace mth
