// Device state of the SVGF denoiser: eight float4 planes per pixel (the reference's Svgf buffers,
// src/denoise/svgf.h:48-58): previous illumination (+history length), previous moments, previous
// normal+depth, previous position+material, current illumination, current moments, two a-trous planes.
#pragma once

#include "internal.h"

namespace phk {

struct SvgfState {
    int w = 0, h = 0;
    bool has_history = false;
    DevBuf<float4> plane[8];
};

void svgf_resize(SvgfState& f, int width, int height);
void svgf_run(SvgfState& f, hipStream_t st, const ph_svgf_input& in, const ph_camera_view& prev_cam, const ph_svgf_settings& s, float4* output);

}  // namespace phk
