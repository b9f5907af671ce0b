// Device state of the SVGF denoiser: eight float4 planes per pixel (the reference's Svgf buffers,
// src/denoise/svgf.h:48-58): previous illumination (+history length), previous moments, previous
// normal+depth, previous position+material, current illumination, current moments, two a-trous planes;
// plus the previous frame's compact geometry when the input is the compact G-buffer.
#pragma once

#include "internal.h"

namespace phk {

struct SvgfState {
    int w = 0, h = 0;
    bool has_history = false;
    DevBuf<float4> plane[8];
    // The previous frame's geometry when the input is the compact G-buffer (planes 2 and 3 hold it for
    // the float4 input); `history_compact` says which format the stored frame is in.
    bool history_compact = false;
    DevBuf<float> prev_depth;
    DevBuf<uint32_t> prev_normal;
    DevBuf<int> prev_material;
};

void svgf_resize(SvgfState& f, int width, int height);
void svgf_run(SvgfState& f, hipStream_t st, const ph_svgf_input& in, const ph_camera_view& prev_cam, const ph_svgf_settings& s, float4* output);
// One frame with the frame-local planes in the library's static device array (no allocation): the
// history is the caller's four planes of width*height float4 (illumination + length, moments,
// normal+depth, position+material), read when `has_history` and rewritten by the frame.
void svgf_run_static(hipStream_t st, int width, int height, const ph_svgf_input& in, const ph_camera_view& prev_cam, const ph_svgf_settings& s,
                     float4* history, bool has_history, float4* output);
// The same denoiser on the compact G-buffer traced from view `cam`.
void svgf_run_compact(SvgfState& f, hipStream_t st, const ph_compact_gbuffer& in, const ph_camera_view& cam, const ph_camera_view& prev_cam,
                      const ph_svgf_settings& s, float4* output);

}  // namespace phk
