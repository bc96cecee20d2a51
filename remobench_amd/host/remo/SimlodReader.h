// The .simlod point file: a 24-byte header (a box of six floats that readers
// ignore) followed by 16-byte point records in remo_point layout, so point i
// starts at byte 24 + 16 i and any range of the cloud is one read.
//
// Plays the role of the reference's raw reader (src/io/RawReader.h:10-31,
// src/io/RawReader.cpp:14-138): the same three calls with the same meaning and
// the same failures -- a file shorter than the header, a payload that is not a
// whole number of records, a file without points, a short read, a range past
// the end -- reported as false plus a message.
#pragma once

#include <cstdint>
#include <string>

#include "remolod.h"

namespace remo {

inline constexpr uint64_t kSimlodHeaderBytes = 24;
inline constexpr uint64_t kSimlodPointBytes = 16;

struct SimlodInfo {
    uint64_t numPoints = 0;
    float headerBox[6] = {};  // as stored; the observed bounds are what the path uses
};

// Record count and header box of `path`.
bool readSimlodHeader(const std::string& path, SimlodInfo& info, std::string* err);

// Observed coordinate bounds {min x, y, z, max x, y, z} over every record, in
// double (the box the header carries is never trusted).
bool readSimlodBounds(const std::string& path, const SimlodInfo& info, double bounds[6], std::string* err);

// Records [first, first + count) into out[0 .. count), each coordinate plus the
// float of `translation` (a float addition, skipped when all three are zero).
bool readSimlodRange(const std::string& path, const SimlodInfo& info, const double translation[3], uint64_t first,
                     uint64_t count, remo_point* out, std::string* err);

}  // namespace remo
