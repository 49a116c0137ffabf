// io.cpp -- velocity models, RTMS shot gathers, raw float arrays, and the
// direct-wave mute, and SEG-Y shot input (the reference's
// src/io/velocity_io.cpp and src/io/seismic_io.cpp).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "rtm/rtm_io.hpp"
#include "rtm_hip.h"

namespace rtm {

namespace {
// Mute one gather: zero before t0 = offset/v + t_pad, raised-cosine ramp over
// max(1, t_taper/dt) samples (src/io/seismic_io.cpp:18-35).
void mute_gather(float sx, float sz, int nrec, const float* rx, const float* rz, int nt, float dt,
                 float* traces, float v_surface, float t_pad, float t_taper) {
    if (v_surface <= 0.0f) return;
    const int ramp = std::max(1, static_cast<int>(t_taper / dt));
    for (int ir = 0; ir < nrec; ++ir) {
        const float ox = rx[ir] - sx;
        const float oz = rz[ir] - sz;
        const float offset = std::sqrt(ox * ox + oz * oz);
        const int i0 = static_cast<int>((offset / v_surface + t_pad) / dt);
        const int i1 = i0 + ramp;
        float* tr = traces + static_cast<std::size_t>(ir) * nt;
        const int stop = std::min(nt, i1);
        for (int it = 0; it < stop; ++it) {
            if (it <= i0) {
                tr[it] = 0.0f;
            } else {
                const float u = static_cast<float>(it - i0) / static_cast<float>(i1 - i0);
                tr[it] *= 0.5f * (1.0f - std::cos(static_cast<float>(M_PI) * u));
            }
        }
    }
}
}  // namespace

void mute_direct_wave(ShotRecord& shot, float v_surface, float t_pad, float t_taper) {
    mute_gather(shot.sx, shot.sz, shot.nrec(), shot.rx.data(), shot.rz.data(), shot.nt, shot.dt,
                shot.traces.data(), v_surface, t_pad, t_taper);
}

namespace io {

// ---- .hdr sidecar: "key = value" lines, '#' comments --------------------------
bool read_header_file(const std::string& path, VelocityHeader& h) {
    std::ifstream f(path);
    if (!f) return false;
    auto trim = [](std::string s) {
        const auto b = s.find_first_not_of(" \t\r\n");
        const auto e = s.find_last_not_of(" \t\r\n");
        return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
    };
    std::string line;
    while (std::getline(f, line)) {
        line = line.substr(0, line.find('#'));
        const auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = trim(line.substr(0, eq));
        const std::string val = trim(line.substr(eq + 1));
        if (key == "nx") h.nx = std::stoi(val);
        else if (key == "nz") h.nz = std::stoi(val);
        else if (key == "dx") h.dx = std::stof(val);
        else if (key == "dz") h.dz = std::stof(val);
        else if (key == "ox") h.ox = std::stof(val);
        else if (key == "oz") h.oz = std::stof(val);
        else if (key == "layout") h.layout = val;
    }
    return true;
}

void write_header_file(const std::string& path, const VelocityHeader& h) {
    std::ofstream f(path);
    if (!f) throw std::runtime_error("cannot write header: " + path);
    f << "# RTM velocity model header\n"
      << "# data file: float32, little-endian, no header\n"
      << "nx = " << h.nx << "\nnz = " << h.nz << "\ndx = " << h.dx << "\ndz = " << h.dz
      << "\nox = " << h.ox << "\noz = " << h.oz << "\nlayout = " << h.layout
      << "   # zfast: v[ix*nz+iz]  |  xfast: v[iz*nx+ix]\n";
}

VelocityModel read_velocity_raw(const std::string& path, const VelocityHeader& h) {
    if (h.nx <= 0 || h.nz <= 0)
        throw std::runtime_error("velocity nx/nz not set (use a .hdr file or --nx/--nz)");
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open velocity file: " + path);
    const auto bytes = static_cast<std::size_t>(f.tellg());
    const std::size_t n = static_cast<std::size_t>(h.nx) * h.nz;
    if (bytes != n * sizeof(float)) {
        std::ostringstream m;
        m << "velocity file size mismatch: " << bytes << " bytes, expected " << n * sizeof(float)
          << " (nx=" << h.nx << ", nz=" << h.nz << ", float32)";
        throw std::runtime_error(m.str());
    }
    std::vector<float> raw(n);
    f.seekg(0);
    f.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(bytes));
    VelocityModel m;
    m.grid.nx = h.nx; m.grid.nz = h.nz; m.grid.dx = h.dx; m.grid.dz = h.dz;
    m.grid.ox = h.ox; m.grid.oz = h.oz;
    if (h.layout == "xfast") {   // v[iz*nx+ix] -> v[ix*nz+iz]
        m.v.resize(n);
        for (int iz = 0; iz < h.nz; ++iz)
            for (int ix = 0; ix < h.nx; ++ix)
                m.v[static_cast<std::size_t>(ix) * h.nz + iz] = raw[static_cast<std::size_t>(iz) * h.nx + ix];
    } else {
        m.v.swap(raw);
    }
    for (float v : m.v)
        if (!(v > 0.0f) || v > 20000.0f)
            throw std::runtime_error("velocity model contains a non-physical value");
    return m;
}

void write_velocity_raw(const std::string& path, const VelocityModel& m) {
    write_raw_floats(path, m.v);
}

// ---- RTMS v1: 32-byte header {"RTMS", 1, nshots, nrec, nt, dt, 0, 0}, then
// per shot sx, sz, rx[nrec], rz[nrec], traces[nrec*nt] (src/io/seismic_io.cpp:39-56)
namespace {
constexpr char kMagic[4] = {'R', 'T', 'M', 'S'};
template <typename T>
void put(std::ofstream& f, const T& v) { f.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <typename T>
void get(std::ifstream& f, T& v) { f.read(reinterpret_cast<char*>(&v), sizeof(T)); }
}  // namespace

void write_shots_raw(const std::string& path, const std::vector<ShotRecord>& shots) {
    if (shots.empty()) throw std::runtime_error("no shots to write");
    const int32_t nrec = shots[0].nrec(), nt = shots[0].nt;
    for (const auto& s : shots)
        if (s.nrec() != nrec || s.nt != nt)
            throw std::runtime_error("RAW format requires constant nrec and nt");
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write shots: " + path);
    f.write(kMagic, 4);
    put(f, int32_t{1});
    put(f, static_cast<int32_t>(shots.size()));
    put(f, nrec);
    put(f, nt);
    put(f, shots[0].dt);
    put(f, int32_t{0});
    put(f, int32_t{0});
    for (const auto& s : shots) {
        put(f, s.sx);
        put(f, s.sz);
        f.write(reinterpret_cast<const char*>(s.rx.data()), static_cast<std::streamsize>(nrec * sizeof(float)));
        f.write(reinterpret_cast<const char*>(s.rz.data()), static_cast<std::streamsize>(nrec * sizeof(float)));
        f.write(reinterpret_cast<const char*>(s.traces.data()),
                static_cast<std::streamsize>(s.traces.size() * sizeof(float)));
    }
    if (!f) throw std::runtime_error("write failed: " + path);
}

std::vector<ShotRecord> read_shots_raw(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open shots: " + path);
    char magic[4] = {};
    f.read(magic, 4);
    if (std::memcmp(magic, kMagic, 4) != 0) throw std::runtime_error(path + ": not an RTMS file (bad magic)");
    int32_t version = 0, nshots = 0, nrec = 0, nt = 0, flags = 0, reserved = 0;
    float dt = 0.0f;
    get(f, version);
    if (version != 1) throw std::runtime_error("unsupported RTMS version");
    get(f, nshots); get(f, nrec); get(f, nt); get(f, dt); get(f, flags); get(f, reserved);
    if (!f || nshots <= 0 || nrec <= 0 || nt <= 0 || !(dt > 0.0f))
        throw std::runtime_error("RTMS header contains invalid values");
    std::vector<ShotRecord> shots(static_cast<std::size_t>(nshots));
    for (auto& s : shots) {
        s.nt = nt;
        s.dt = dt;
        s.rx.resize(nrec);
        s.rz.resize(nrec);
        s.traces.resize(static_cast<std::size_t>(nrec) * nt);
        get(f, s.sx);
        get(f, s.sz);
        f.read(reinterpret_cast<char*>(s.rx.data()), static_cast<std::streamsize>(nrec * sizeof(float)));
        f.read(reinterpret_cast<char*>(s.rz.data()), static_cast<std::streamsize>(nrec * sizeof(float)));
        f.read(reinterpret_cast<char*>(s.traces.data()),
               static_cast<std::streamsize>(s.traces.size() * sizeof(float)));
        if (!f) throw std::runtime_error("RTMS file truncated: " + path);
    }
    return shots;
}

// ---- SEG-Y rev 0/1 input (src/io/seismic_io.cpp:126-249) -------------------
// Layout: 3200-byte text header, 400-byte binary header, then traces of a
// 240-byte header and ns samples, everything big-endian.  Read: the sample
// interval (µs), samples per trace and sample format from the binary header
// (file bytes 3217, 3221, 3225, 1-based); per trace the field record number
// (bytes 9-12), the coordinate scalar (71-72), source x (73-76) and group x
// (81-84).  Consecutive traces with the same record number form one shot; with
// record number 0 a change of source x starts a new shot.  Depths are not read
// from the file: they are the caller's src_depth and rec_depth.
namespace {

int16_t load_be16(const unsigned char* p) { return static_cast<int16_t>((p[0] << 8) | p[1]); }
uint32_t load_be32u(const unsigned char* p) {
    return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
}
int32_t load_be32(const unsigned char* p) { return static_cast<int32_t>(load_be32u(p)); }

// IBM System/360 single precision: sign, 7-bit base-16 exponent biased by 64,
// 24-bit fraction F with value 0.F * 16^(e-64).  The value is exact in double
// (|exponent| <= 2^258), so it is formed exactly and rounded once to float.
float ibm32_to_float(uint32_t w) {
    const uint32_t frac = w & 0x00ffffffu;
    const int e16 = static_cast<int>((w >> 24) & 0x7fu) - 64;
    const double mag = std::ldexp(static_cast<double>(frac), 4 * e16 - 24);
    return static_cast<float>((w >> 31) ? -mag : mag);
}

// Coordinate scalar: > 0 multiplies, < 0 divides by its magnitude, 0 leaves as is.
float scaled_coordinate(int32_t c, int16_t scalar) {
    if (scalar > 0) return static_cast<float>(c) * static_cast<float>(scalar);
    if (scalar < 0) return static_cast<float>(c) / static_cast<float>(-scalar);
    return static_cast<float>(c);
}

int segy_sample_bytes(int format) {
    switch (format) {
        case 1: case 2: case 5: return 4;
        case 3: return 2;
        default: return 0;
    }
}

void decode_segy_samples(int format, const unsigned char* p, int ns, float* out) {
    switch (format) {
        case 1:
            for (int k = 0; k < ns; ++k) out[k] = ibm32_to_float(load_be32u(p + 4 * k));
            break;
        case 5:
            for (int k = 0; k < ns; ++k) {
                const uint32_t bits = load_be32u(p + 4 * k);
                std::memcpy(out + k, &bits, 4);
            }
            break;
        case 2:
            for (int k = 0; k < ns; ++k) out[k] = static_cast<float>(load_be32(p + 4 * k));
            break;
        case 3:
            for (int k = 0; k < ns; ++k) out[k] = static_cast<float>(load_be16(p + 2 * k));
            break;
    }
}

}  // namespace

std::vector<ShotRecord> read_shots_segy(const std::string& path, float src_depth, float rec_depth) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open SEG-Y: " + path);
    const long long file_bytes = static_cast<long long>(f.tellg());
    if (file_bytes < 3600) throw std::runtime_error("SEG-Y too short: " + path);
    unsigned char bin[400];
    f.seekg(3200);
    f.read(reinterpret_cast<char*>(bin), sizeof bin);
    const int dt_us = load_be16(bin + 16);
    const int ns = load_be16(bin + 20);
    const int format = load_be16(bin + 24);
    if (ns <= 0) throw std::runtime_error("SEG-Y: invalid samples per trace");
    if (dt_us <= 0) throw std::runtime_error("SEG-Y: invalid sample interval");
    const int bps = segy_sample_bytes(format);
    if (bps == 0) {
        std::ostringstream m;
        m << "SEG-Y sample format " << format << " is not supported (use 1, 2, 3 or 5)";
        throw std::runtime_error(m.str());
    }
    const long long trace_bytes = 240 + static_cast<long long>(ns) * bps;
    const long long ntraces = (file_bytes - 3600) / trace_bytes;
    if (ntraces <= 0) throw std::runtime_error("SEG-Y contains no traces");

    const float dt = static_cast<float>(dt_us) * 1e-6f;
    std::vector<ShotRecord> shots;
    std::vector<unsigned char> tr(static_cast<std::size_t>(trace_bytes));
    int32_t shot_record = 0;
    float shot_x = 0.0f;
    f.seekg(3600);
    for (long long t = 0; t < ntraces; ++t) {
        if (!f.read(reinterpret_cast<char*>(tr.data()), static_cast<std::streamsize>(trace_bytes))) break;
        const int32_t record = load_be32(tr.data() + 8);
        const int16_t scalar = load_be16(tr.data() + 70);
        const float sx = scaled_coordinate(load_be32(tr.data() + 72), scalar);
        const float gx = scaled_coordinate(load_be32(tr.data() + 80), scalar);
        if (shots.empty() || record != shot_record || (record == 0 && sx != shot_x)) {
            ShotRecord s;
            s.sx = sx;
            s.sz = src_depth;
            s.nt = ns;
            s.dt = dt;
            shots.push_back(std::move(s));
            shot_record = record;
            shot_x = sx;
        }
        ShotRecord& s = shots.back();
        s.rx.push_back(gx);
        s.rz.push_back(rec_depth);
        const std::size_t at = s.traces.size();
        s.traces.resize(at + static_cast<std::size_t>(ns));
        decode_segy_samples(format, tr.data() + 240, ns, s.traces.data() + at);
    }
    if (shots.empty()) throw std::runtime_error("SEG-Y: no shots decoded");
    return shots;
}

std::vector<ShotRecord> read_shots(const std::string& path, float src_depth, float rec_depth) {
    std::string ext;
    const auto dot = path.find_last_of('.');
    if (dot != std::string::npos) ext = path.substr(dot);
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    if (ext == ".sgy" || ext == ".segy") return read_shots_segy(path, src_depth, rec_depth);
    return read_shots_raw(path);
}

void write_raw_floats(const std::string& path, const std::vector<float>& v) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write: " + path);
    f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(float)));
}

std::vector<float> read_raw_floats(const std::string& path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open: " + path);
    const auto bytes = static_cast<std::size_t>(f.tellg());
    if (bytes % sizeof(float) != 0) throw std::runtime_error("file is not a float32 array: " + path);
    std::vector<float> v(bytes / sizeof(float));
    f.seekg(0);
    f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(bytes));
    return v;
}

}  // namespace io
}  // namespace rtm

extern "C" void rtm_hip_mute_direct_wave(const rtm_hip_shot* shot, float* traces, float v_surface,
                                         float t_pad, float t_taper) {
    if (!shot || !traces) return;
    rtm::mute_gather(shot->sx, shot->sz, shot->nrec, shot->rx, shot->rz, shot->nt, shot->dt, traces,
                     v_surface, t_pad, t_taper);
}
