// rtm_engine.hpp -- the engine abstraction (same surface as the reference's
// include/rtm_engine.hpp:1-90): every engine implements setup, the two
// propagations and imaging; migrate and model_shot have default drivers.
#pragma once
#include <string>
#include <vector>
#include "rtm_types.hpp"

namespace rtm {

// Wall-clock seconds per stage.
struct StageTimes {
    double io = 0, forward = 0, backward = 0, imaging = 0, total = 0;
    double h2d = 0, kernel = 0, d2h = 0;
    void reset() { *this = StageTimes(); }
};

class RTMEngine {
public:
    virtual ~RTMEngine() = default;
    virtual const char* name() const = 0;

    virtual void setup(const VelocityModel& model, const RTMParams& par, const TimeAxis& ta) = 0;

    // snapshots: if non-null, receives the interior wavefield every
    // store_interval steps (num_snapshots()*nx*nz); recorded: if non-null,
    // receives modelled traces (nrec*nt, trace-major).
    virtual void forward_propagation(const ShotRecord& shot, std::vector<float>* snapshots,
                                     std::vector<float>* recorded) = 0;

    // Back-propagates the recorded data and accumulates the image.
    virtual void backward_propagation(const ShotRecord& shot, const std::vector<float>& snapshots,
                                      std::vector<float>& image,
                                      std::vector<float>& illumination) = 0;

    // Zero-lag cross-correlation of one time step (interior snapshot,
    // extended receiver field).
    virtual void imaging(const float* fwd_snapshot, const float* bwd_extended,
                         std::vector<float>& image, std::vector<float>& illumination) = 0;

    virtual void migrate(const std::vector<ShotRecord>& shots, std::vector<float>& image,
                         std::vector<float>& illumination);
    virtual void model_shot(ShotRecord& shot);

    const Grid&      grid() const { return g_; }
    const RTMParams& params() const { return par_; }
    const TimeAxis&  time_axis() const { return ta_; }
    int              num_snapshots() const { return nsnap_; }
    std::size_t      snapshot_bytes() const {
        return static_cast<std::size_t>(nsnap_) * g_.n_interior() * sizeof(float);
    }

    virtual int         num_threads() const { return 1; }
    virtual int         num_devices() const { return 0; }
    virtual std::string device_name() const { return "cpu"; }
    virtual bool        is_gpu() const { return false; }

    StageTimes times;

protected:
    Grid       g_;
    RTMParams  par_;
    TimeAxis   ta_;
    int        nsnap_ = 0;
    std::vector<float> snap_;
};

// Image post-processing (applied once after all shots).
void illumination_compensation(std::vector<float>& image, const std::vector<float>& illumination,
                               float eps_relative);
void laplacian_filter(std::vector<float>& image, const Grid& g);

}  // namespace rtm
