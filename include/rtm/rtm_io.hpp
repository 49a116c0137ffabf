// rtm_io.hpp -- velocity and shot-gather I/O and the direct-wave mute (same
// surface as the reference's include/rtm_io.hpp:9-37).
#pragma once
#include <string>
#include <vector>
#include "rtm_types.hpp"

namespace rtm {

// Zeroes each trace before offset/v_surface + t_pad, then a raised-cosine
// taper of length t_taper.
void mute_direct_wave(ShotRecord& shot, float v_surface, float t_pad, float t_taper);

namespace io {

struct VelocityHeader {
    int   nx = 0, nz = 0;
    float dx = 10.0f, dz = 10.0f, ox = 0.0f, oz = 0.0f;
    std::string layout = "zfast";   // zfast: v[ix*nz+iz]; xfast: v[iz*nx+ix]
};
bool read_header_file(const std::string& path, VelocityHeader& h);
void write_header_file(const std::string& path, const VelocityHeader& h);

VelocityModel read_velocity_raw(const std::string& path, const VelocityHeader& h);
void          write_velocity_raw(const std::string& path, const VelocityModel& m);

// RTMS v1 little-endian shot container.
void                    write_shots_raw(const std::string& path, const std::vector<ShotRecord>& shots);
std::vector<ShotRecord> read_shots_raw(const std::string& path);
// SEG-Y rev 0/1 (constant trace length; sample formats 1 IBM float, 2 int32,
// 3 int16, 5 IEEE float): traces grouped into shots by field record number,
// source and receiver x from the trace headers with the coordinate scalar,
// depths from the arguments (src/io/seismic_io.cpp:126-249).
std::vector<ShotRecord> read_shots_segy(const std::string& path, float src_depth, float rec_depth);
// Dispatch on extension: .sgy/.segy -> SEG-Y, anything else -> RTMS.
std::vector<ShotRecord> read_shots(const std::string& path, float src_depth, float rec_depth);

void               write_raw_floats(const std::string& path, const std::vector<float>& v);
std::vector<float> read_raw_floats(const std::string& path);

}  // namespace io
}  // namespace rtm
