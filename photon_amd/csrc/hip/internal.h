// Host-side internals of the device library: error handling, owned device buffers, and the launchers
// each translation unit exports (kernels are launched only from the unit that defines them).
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "dpath.h"
#include "photon_hip.h"

namespace phk {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define PH_CHECK(x)                                                                                               \
    do {                                                                                                          \
        hipError_t e_ = (x);                                                                                      \
        if (e_ != hipSuccess) throw ::phk::HipError(std::string("HIP error: ") + hipGetErrorString(e_) + " (" #x ")"); \
    } while (0)

// The calling thread's last error message (ph_last_error) and the entry-point wrapper that turns
// exceptions into a negative status. A failed HIP call also leaves HIP's last-error state set; it is
// cleared so that a later call does not report it again.
int set_error(const std::string& msg);

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        (void)hipGetLastError();
        return set_error("out of host memory");
    } catch (const std::exception& e) {
        (void)hipGetLastError();
        return set_error(e.what());
    }
}

// An owning device allocation.
template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void allocate(size_t n) {
        if (n == count && ptr) return;
        release();
        if (n == 0) return;
        PH_CHECK(hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T)));
        count = n;
    }
    void upload(const void* src, size_t n, hipStream_t st) {
        allocate(n);
        if (n) PH_CHECK(hipMemcpyAsync(ptr, src, n * sizeof(T), hipMemcpyHostToDevice, st));
    }
    void zero(hipStream_t st) {
        if (ptr) PH_CHECK(hipMemsetAsync(ptr, 0, count * sizeof(T), st));
    }
    size_t bytes() const { return count * sizeof(T); }
};

template <class T>
const T* or_null(const DevBuf<T>& b) {
    return b.count ? b.ptr : nullptr;
}

// ---- render.hip -------------------------------------------------------------------------------------
struct DisplayParams {
    int w, h, mode;
    float exposure;
    int tonemap;
    float inv_samples, samples, max_depth, heat_scale;
    int denoised;
};

void launch_render(hipStream_t st, const phd::DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame0, int spp, uint32_t seed,
                   const phd::PathOpts& opt, float4* accum, float4* albedo, float4* nd);
void launch_sample(hipStream_t st, const phd::DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame, uint32_t seed, const phd::PathOpts& opt,
                   float4* color, float4* albedo, float4* nd, float4* position, float4* emission);
// The device BVH build behind ph_bvh_build: writes the flat nodes (room for max(1, 2n - 1)) and the
// triangle order (n entries) to host memory and returns the node count.
int64_t bvh_build_device(const float* positions, int num_vertices, const int* indices, int n, int bins, int max_leaf, float traversal_cost,
                         ph_bvh_node* nodes_out, int* order_out, ph_bvh_build_stats& stats);

// Guide tables of `rows` CDFs of n + 1 entries each (row r at cdf + r * stride): guide[r * (m + 1) + k]
// is the index the CDF search returns for u = k / m (m a power of two).
void build_cdf_guides(hipStream_t st, const float* cdf, int rows, int n, size_t stride, int m, int* guide);

// The same sample written as the compact G-buffer (photon_hip.h, ph_compact_gbuffer).
void launch_sample_compact(hipStream_t st, const phd::DScene& s, const ph_camera_view& cam, int w, int h, uint32_t frame, uint32_t seed,
                           const phd::PathOpts& opt, const ph_compact_gbuffer& out);
void launch_heatmap(hipStream_t st, const phd::DScene& s, const ph_camera_view& cam, int w, int h, float* out);
void launch_display(hipStream_t st, const DisplayParams& p, const float4* accum, const float4* denoised, const float4* albedo, const float4* nd,
                    const float4* variance, const float* heat, uint32_t* out);
// Returns false when (block, depth) is not an instantiated shape.
// Persistent launches stage the top `lds_nodes` BVH nodes in each block's LDS.
bool launch_trace(hipStream_t st, const phd::DScene& s, const ph_ray* rays, int n, bool anyhit, ph_hit* hits, uint32_t* counts, int block, int depth,
                  bool persistent, int lds_nodes, int grid_blocks);

}  // namespace phk
