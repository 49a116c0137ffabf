// rtm_factory.hpp -- engine selection by name (same surface as the
// reference's include/rtm_factory.hpp:19-23).
//
//   hip          HipRTM, one shot at a time on one MI355X (gfx950)
//
// Throws std::runtime_error for an unknown name.
#pragma once
#include <memory>
#include <string>
#include <vector>
#include "rtm_engine.hpp"

namespace rtm {

std::unique_ptr<RTMEngine> make_engine(const std::string& engine_name, int requested_gpus = 1);
std::vector<std::string> list_engines();

}  // namespace rtm
