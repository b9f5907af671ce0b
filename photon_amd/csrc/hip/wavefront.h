// The wavefront path tracing engine: paths in flight as structure-of-arrays state in HBM, advanced
// one bounce at a time by persistent traversal and shading kernels that consume compacted queues.
#pragma once

#include <memory>

#include "internal.h"

namespace phk {

// One progressive render request: `spp` samples per pixel with frame indices frame0.., added to
// `accum` (rgb sum, w sum of luminance^2) with the G-buffer of the first sample in albedo / nd.
struct WavefrontJob {
    phd::DScene scene;
    ph_camera_view camera;
    int width, height, spp;
    uint32_t frame0, seed;
    phd::PathOpts opt;
    float4* accum;
    float4* albedo;
    float4* nd;
};

class WavefrontEngine {
public:
    WavefrontEngine();
    ~WavefrontEngine();
    WavefrontEngine(const WavefrontEngine&) = delete;
    WavefrontEngine& operator=(const WavefrontEngine&) = delete;

    void set_config(const ph_wavefront_config& cfg);
    ph_wavefront_config config() const;
    void render(hipStream_t st, const WavefrontJob& job, bool counting, ph_render_stats& stats);
    ph_wavefront_config autotune(hipStream_t st, const WavefrontJob& job);

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
};

}  // namespace phk
