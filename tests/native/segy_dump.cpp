// segy_dump -- reads a shot file through rtm::io::read_shots (the extension
// picks SEG-Y or RTMS) and writes what it decoded as plain native-endian
// records, so that two builds of the reader can be compared byte for byte:
//   int32 nshots, then per shot: int32 nt, float dt, float sx, float sz,
//   int32 nrec, float rx[nrec], float rz[nrec], float traces[nrec*nt]
// On an exception it prints "error: <message>" and exits with status 3.
//
// Compiled against this repository's include/rtm and io.cpp, and by
// oracle/Makefile against the reference's include/ and its compiled sources:
// both declare the same rtm::io::read_shots and ShotRecord.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <vector>

#include "rtm_io.hpp"

namespace {
template <typename T>
void put(std::FILE* f, const T& v) { std::fwrite(&v, sizeof(T), 1, f); }
void put_floats(std::FILE* f, const std::vector<float>& v) { std::fwrite(v.data(), sizeof(float), v.size(), f); }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: segy_dump SHOTS SRC_DEPTH REC_DEPTH OUT\n");
        return 2;
    }
    std::vector<rtm::ShotRecord> shots;
    try {
        shots = rtm::io::read_shots(argv[1], std::strtof(argv[2], nullptr), std::strtof(argv[3], nullptr));
    } catch (const std::exception& e) {
        std::printf("error: %s\n", e.what());
        return 3;
    }
    std::FILE* f = std::fopen(argv[4], "wb");
    if (!f) return 2;
    put(f, static_cast<int32_t>(shots.size()));
    for (const auto& s : shots) {
        put(f, static_cast<int32_t>(s.nt));
        put(f, s.dt);
        put(f, s.sx);
        put(f, s.sz);
        put(f, static_cast<int32_t>(s.rx.size()));
        put_floats(f, s.rx);
        put_floats(f, s.rz);
        put_floats(f, s.traces);
    }
    return std::fclose(f) == 0 ? 0 : 2;
}
