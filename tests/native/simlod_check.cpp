// Test driver for the C++ .simlod reader (no device): prints what the reader
// returns for a file as one JSON line, or its error message.
//   simlod_check FILE [FIRST COUNT]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "remo/SimlodReader.h"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string path = argv[1];
    std::string err;
    remo::SimlodInfo info;
    if (!remo::readSimlodHeader(path, info, &err)) {
        std::printf("{\"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    double b[6];
    if (!remo::readSimlodBounds(path, info, b, &err)) {
        std::printf("{\"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    const double t[3] = {-b[0], -b[1], -b[2]};
    const uint64_t first = argc > 3 ? std::strtoull(argv[2], nullptr, 10) : 0;
    const uint64_t count = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : info.numPoints;
    std::vector<remo_point> pts(count);
    if (!remo::readSimlodRange(path, info, t, first, count, pts.data(), &err)) {
        std::printf("{\"error\": \"%s\"}\n", err.c_str());
        return 1;
    }
    unsigned long long h = 1469598103934665603ull;  // FNV-1a over the translated records
    const auto* bytes = reinterpret_cast<const unsigned char*>(pts.data());
    for (size_t i = 0; i < pts.size() * sizeof(remo_point); ++i) h = (h ^ bytes[i]) * 1099511628211ull;
    std::printf("{\"num_points\": %llu, \"header_box\": [%.9g, %.9g, %.9g, %.9g, %.9g, %.9g], "
                "\"bounds\": [%.17g, %.17g, %.17g, %.17g, %.17g, %.17g], \"box_size\": [%.9g, %.9g, %.9g], "
                "\"fnv\": \"%016llx\"}\n",
                (unsigned long long)info.numPoints, info.headerBox[0], info.headerBox[1], info.headerBox[2],
                info.headerBox[3], info.headerBox[4], info.headerBox[5], b[0], b[1], b[2], b[3], b[4], b[5],
                double(float(b[3] + t[0])), double(float(b[4] + t[1])), double(float(b[5] + t[2])), h);
    return 0;
}
