// .simlod reading with positioned reads (no stream state, so a range read needs
// no seek and concurrent readers of one file do not interfere).
#include "remo/SimlodReader.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <vector>

namespace remo {

namespace {

bool fail(std::string* err, const std::string& what) {
    if (err) *err = what;
    return false;
}

// Closes the descriptor on every way out.
struct Fd {
    int fd = -1;
    explicit Fd(const std::string& path) : fd(::open(path.c_str(), O_RDONLY | O_CLOEXEC)) {}
    ~Fd() {
        if (fd >= 0) ::close(fd);
    }
};

// Exactly `bytes` bytes at `offset`, or false (a short file, an I/O error).
bool readFully(int fd, uint64_t offset, void* dst, uint64_t bytes) {
    auto* p = static_cast<char*>(dst);
    while (bytes > 0) {
        const ssize_t got = ::pread(fd, p, size_t(std::min<uint64_t>(bytes, 1ull << 30)), off_t(offset));
        if (got < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        if (got == 0) return false;
        p += got;
        offset += uint64_t(got);
        bytes -= uint64_t(got);
    }
    return true;
}

// Records per read while scanning for the bounds: 1 MiB of records.
constexpr uint64_t kScanRecords = 65536;

}  // namespace

bool readSimlodHeader(const std::string& path, SimlodInfo& info, std::string* err) {
    Fd f(path);
    if (f.fd < 0) return fail(err, "cannot open " + path);
    struct stat st{};
    if (::fstat(f.fd, &st) != 0) return fail(err, "cannot stat " + path);
    const uint64_t size = uint64_t(st.st_size);
    if (size < kSimlodHeaderBytes) return fail(err, "file is smaller than a .simlod header: " + path);
    if ((size - kSimlodHeaderBytes) % kSimlodPointBytes != 0)
        return fail(err, "point data is not a whole number of 16-byte records (truncated?): " + path);
    if (!readFully(f.fd, 0, info.headerBox, sizeof(info.headerBox)))
        return fail(err, "failed reading the .simlod header of " + path);
    info.numPoints = (size - kSimlodHeaderBytes) / kSimlodPointBytes;
    if (info.numPoints == 0) return fail(err, ".simlod file holds no points: " + path);
    return true;
}

bool readSimlodBounds(const std::string& path, const SimlodInfo& info, double bounds[6], std::string* err) {
    Fd f(path);
    if (f.fd < 0) return fail(err, "cannot open " + path);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    std::vector<remo_point> buf(size_t(std::min<uint64_t>(kScanRecords, info.numPoints)));
    for (uint64_t at = 0; at < info.numPoints;) {
        const uint64_t n = std::min<uint64_t>(kScanRecords, info.numPoints - at);
        if (!readFully(f.fd, kSimlodHeaderBytes + at * kSimlodPointBytes, buf.data(), n * kSimlodPointBytes))
            return fail(err, "short read on " + path);
        for (uint64_t i = 0; i < n; ++i) {
            const double c[3] = {buf[i].x, buf[i].y, buf[i].z};
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], c[k]);
                hi[k] = std::max(hi[k], c[k]);
            }
        }
        at += n;
    }
    for (int k = 0; k < 3; ++k) {
        bounds[k] = lo[k];
        bounds[3 + k] = hi[k];
    }
    return true;
}

bool readSimlodRange(const std::string& path, const SimlodInfo& info, const double translation[3], uint64_t first,
                     uint64_t count, remo_point* out, std::string* err) {
    if (first > info.numPoints || count > info.numPoints - first)
        return fail(err, ".simlod point range runs past the end of the file: " + path);
    Fd f(path);
    if (f.fd < 0) return fail(err, "cannot open " + path);
    if (!readFully(f.fd, kSimlodHeaderBytes + first * kSimlodPointBytes, out, count * kSimlodPointBytes))
        return fail(err, "short read on " + path);
    const float t[3] = {float(translation[0]), float(translation[1]), float(translation[2])};
    if (t[0] == 0.0f && t[1] == 0.0f && t[2] == 0.0f) return true;
    for (uint64_t i = 0; i < count; ++i) {
        out[i].x += t[0];
        out[i].y += t[1];
        out[i].z += t[2];
    }
    return true;
}

}  // namespace remo
