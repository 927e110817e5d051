// pdbio.cpp -- see pdbio.hpp.
#include "pdbio.hpp"

#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <sstream>

namespace mth {

namespace {

// Column field [start, start + len) of a record, "" past its end.
std::string col(const std::string &line, size_t start, size_t len) {
    if (start >= line.size()) return "";
    return line.substr(start, len);
}

// Chunks as fgets(buffer, 80) returns them: at most 79 characters, the
// newline kept with the chunk it ends.
std::vector<std::string> fgets_chunks(const std::string &text) {
    std::vector<std::string> out;
    size_t pos = 0;
    while (pos < text.size()) {
        size_t nl = text.find('\n', pos);
        const size_t end = nl == std::string::npos ? text.size() : nl + 1;
        size_t take = std::min<size_t>(79, end - pos);
        out.push_back(text.substr(pos, take));
        pos += take;
    }
    return out;
}

std::string first_token_space(const std::string &s) {
    const size_t b = s.find_first_not_of(' ');
    if (b == std::string::npos) return "";
    const size_t e = s.find(' ', b);
    return s.substr(b, e == std::string::npos ? std::string::npos : e - b);
}

}  // namespace

std::vector<PdbAtom> read_pdb(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw IoError(path + ": cannot open");
    std::stringstream ss;
    ss << in.rdbuf();
    const std::vector<std::string> chunks = fgets_chunks(ss.str());
    size_t counted = 0;
    for (const std::string &c : chunks)
        if (c.compare(0, 4, "ATOM") == 0) counted++;
    std::vector<PdbAtom> atoms;
    atoms.reserve(counted);
    for (const std::string &c : chunks) {
        if (first_token_space(c) != "ATOM") continue;
        PdbAtom a;
        a.id = std::atoi(col(c, 6, 5).c_str());
        a.name = first_token_space(col(c, 12, 4));
        if (a.name.empty()) throw IoError(path + ": ATOM record without an atom name");
        a.altLoc = c.size() > 16 ? c[16] : '\0';
        a.resName = col(c, 17, 3);
        a.chain = c.size() > 21 ? c[21] : '\0';
        a.resid = std::atoi(col(c, 22, 4).c_str());
        a.x = std::atof(col(c, 30, 8).c_str());
        a.y = std::atof(col(c, 38, 8).c_str());
        a.z = std::atof(col(c, 46, 8).c_str());
        a.occupancy = std::atof(col(c, 54, 6).c_str());
        a.beta = std::atof(col(c, 60, 6).c_str());
        atoms.push_back(a);
    }
    if (atoms.size() != counted)
        throw IoError(path + ": records starting with 'ATOM' that are not ATOM records (e.g. serials >= 100000)");
    return atoms;
}

void write_pdb(const std::string &path, const std::vector<PdbAtom> &atoms) {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw IoError(path + ": cannot write");
    const bool hex = atoms.size() >= 100000;
    for (size_t i = 0; i < atoms.size(); i++) {
        const PdbAtom &a = atoms[i];
        std::fprintf(f, hex ? "ATOM  %5x %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n"
                            : "ATOM  %5d %-4s%c%3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f\n",
                     (int)(i + 1), a.name.c_str(), a.altLoc, a.resName.c_str(), a.chain, a.resid, a.x, a.y, a.z,
                     a.occupancy, a.beta);
    }
    std::fprintf(f, "END");
    std::fclose(f);
}

DcdWriter::DcdWriter(int natoms, int nframes, int first_frame, float timestep, int save_freq)
    : n_(natoms), nfile_(nframes), npriv_(first_frame), nsavc_(save_freq), delta_(timestep) {
    remark1_ = "REMARKS CREATED BY dcdio.c";
    std::time_t now = std::time(nullptr);
    remark2_ = std::string("REMARKS DATE: ") + std::asctime(std::localtime(&now));
}

void DcdWriter::create(const std::string &path) const {
    FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw IoError(path + ": cannot write");
    auto i32 = [f](int32_t v) { std::fwrite(&v, 4, 1, f); };
    i32(84);
    std::fwrite("CORD", 4, 1, f);
    i32(nfile_);
    i32(npriv_);
    i32(nsavc_);
    i32(npriv_ - nsavc_);
    for (int k = 0; k < 5; k++) i32(0);
    std::fwrite(&delta_, 4, 1, f);
    i32(0);   // no unit cell
    for (int k = 0; k < 8; k++) i32(0);
    i32(24);
    i32(84);
    i32(164);
    i32(2);
    for (const std::string *r : {&remark1_, &remark2_}) {
        // NUL-padded to 80 bytes; longer remarks cut to 79 characters
        char title[80];
        std::memset(title, 0, sizeof title);
        std::memcpy(title, r->data(), r->size() > 80 ? 79 : r->size());
        std::fwrite(title, 80, 1, f);
    }
    i32(164);
    i32(4);
    i32(n_);
    i32(4);
    std::fclose(f);
}

void DcdWriter::append(const std::string &path, const float *x, const float *y, const float *z) const {
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f) throw IoError(path + ": cannot append");
    const int32_t bytes = 4 * n_;
    for (const float *c : {x, y, z}) {
        std::fwrite(&bytes, 4, 1, f);
        std::fwrite(c, 4, (size_t)n_, f);
        std::fwrite(&bytes, 4, 1, f);
    }
    std::fclose(f);
}

}  // namespace mth
