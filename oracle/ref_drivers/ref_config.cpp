// Driver (this repo's own code) around the reference's config reader,
// src/configreader.cpp + src/wrapper.cpp, compiled where they lie.
// usage: ref_config FILE [name=value ...] < queries
// Each query line: "<kind> <name> [default]" with kind in
//   str | int | long | float | yesno | masked
// A query without a default uses the reference's "required" form (DIE on a
// missing key).  Prints one "name<TAB>value" line per query on stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "configreader.h"

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: ref_config FILE [name=value ...]\n"); return 2; }
    // parseParametersFile reads overrides from argv[2..]
    parseParametersFile(argv[1], argc, argv);
    char line[4096];
    while (std::fgets(line, sizeof line, stdin)) {
        char kind[64] = {0}, name[256] = {0}, def[2048] = {0};
        int n = std::sscanf(line, "%63s %255s %2047[^\n]", kind, name, def);
        if (n < 2) continue;
        bool has_def = n == 3;
        char out[4096];
        if (!std::strcmp(kind, "str")) {
            if (has_def) getParameter(out, name, def); else getParameter(out, name);
        } else if (!std::strcmp(kind, "masked")) {
            if (has_def) getMaskedParameter(out, name, def); else getMaskedParameter(out, name);
        } else if (!std::strcmp(kind, "int")) {
            int v = has_def ? getIntegerParameter(name, std::atoi(def)) : getIntegerParameter(name);
            std::sprintf(out, "%d", v);
        } else if (!std::strcmp(kind, "long")) {
            long long v = has_def ? getLongIntegerParameter(name, std::atol(def)) : getLongIntegerParameter(name);
            std::sprintf(out, "%lld", v);
        } else if (!std::strcmp(kind, "float")) {
            float v = has_def ? getFloatParameter(name, (float)std::atof(def)) : getFloatParameter(name);
            std::sprintf(out, "%.9g", v);
        } else if (!std::strcmp(kind, "yesno")) {
            int v = has_def ? getYesNoParameter(name, std::atoi(def)) : getYesNoParameter(name);
            std::sprintf(out, "%d", v);
        } else {
            continue;
        }
        std::fprintf(stderr, "@@%s\t%s\n", name, out);
    }
    return 0;
}
