// A scene resident in HBM: owns one allocation per flat array and the device view (DScene) over them.
// Used by the renderer (one scene) and by the serving loop (several scenes resident at once).
#pragma once

#include <cmath>
#include <cstring>
#include <deque>
#include <vector>

#include "internal.h"

namespace phk {

// Node relabelling for the device copy of the BVH: the first `top` nodes in breadth-first order from the
// root come first, the others follow in their original order (so each deeper subtree stays contiguous).
// Child links are renumbered; leaves keep their triangle ranges. The tree, the traversal order and every
// hit are unchanged; the top levels, which every ray visits, form a prefix that kernels can stage in LDS.
inline std::vector<ph_bvh_node> relabel_top_first(const ph_bvh_node* nodes, int n, int top, int& top_count) {
    std::vector<int> order;
    order.reserve(static_cast<size_t>(n));
    std::vector<int> index(static_cast<size_t>(n), -1);
    std::deque<int> bfs{0};
    while (!bfs.empty() && int(order.size()) < top) {
        const int i = bfs.front();
        bfs.pop_front();
        if (i < 0 || i >= n || index[size_t(i)] >= 0) continue;
        index[size_t(i)] = int(order.size());
        order.push_back(i);
        const ph_bvh_node& nd = nodes[i];
        if (nd.lcount == 0 && nd.left >= 0) bfs.push_back(nd.left);
        if (nd.rcount == 0 && nd.right >= 0) bfs.push_back(nd.right);
    }
    top_count = int(order.size());
    for (int i = 0; i < n; i++)
        if (index[size_t(i)] < 0) {
            index[size_t(i)] = int(order.size());
            order.push_back(i);
        }
    std::vector<ph_bvh_node> out(static_cast<size_t>(n));
    for (int k = 0; k < n; k++) {
        ph_bvh_node nd = nodes[order[size_t(k)]];
        if (nd.lcount == 0 && nd.left >= 0) nd.left = index[size_t(nd.left)];
        if (nd.rcount == 0 && nd.right >= 0) nd.right = index[size_t(nd.right)];
        out[size_t(k)] = nd;
    }
    return out;
}

class DeviceScene {
public:
    // Nodes relabelled to the front of the device BVH (at most this many can be staged in LDS).
    static constexpr int kTopNodes = 4096;

    const phd::DScene& view() const { return view_; }
    float extent() const { return extent_; }
    bool ready() const { return ready_; }
    size_t bytes() const {
        size_t b = 0;
        b += nodes_.bytes() + tris_.bytes() + materials_.bytes() + indices_.bytes() + tri_emissive_.bytes();
        b += normals_.bytes() + uvs_.bytes() + tangents_.bytes() + srgb_lut_.bytes() + env_texels_.bytes();
        b += env_marginal_.bytes() + env_conditional_.bytes() + textures_.bytes() + texels_.bytes();
        b += guide_marginal_.bytes() + guide_conditional_.bytes();
        b += emissive_.bytes() + alias_.bytes() + punctual_.bytes();
        return b;
    }

    static void validate(const ph_scene_desc* d) {
        if (!d) throw std::invalid_argument("null scene descriptor");
        if (!d->nodes || d->num_nodes <= 0) throw std::invalid_argument("the scene is not built (no BVH nodes)");
        if (d->num_tris < 0 || (d->num_tris > 0 && (!d->tris || !d->indices))) throw std::invalid_argument("missing triangle arrays");
        if (!d->materials || d->num_materials <= 0) throw std::invalid_argument("the scene has no materials");
        if (!d->srgb_lut) throw std::invalid_argument("missing sRGB table");
        if (d->num_emissive > 0 && (!d->emissive || !d->emissive_alias || !d->tri_emissive)) throw std::invalid_argument("missing light tables");
        if (d->env_texels && (d->env_width <= 0 || d->env_height <= 0 || !d->env_marginal || !d->env_conditional))
            throw std::invalid_argument("incomplete environment image");
        if (d->num_textures > 0 && (!d->textures || !d->texels)) throw std::invalid_argument("missing texture pool");
    }

    // Copies every array (GpuRenderer::upload, gpu_renderer.cu:243-297). Synchronous.
    void upload(const ph_scene_desc& d, hipStream_t st) {
        ready_ = false;
        int top = 0;
        const std::vector<ph_bvh_node> relabelled = relabel_top_first(d.nodes, d.num_nodes, kTopNodes, top);
        nodes_.upload(relabelled.data(), size_t(d.num_nodes) * 4, st);
        PH_CHECK(hipStreamSynchronize(st));  // the host copy above goes out of scope
        indices_.upload(d.indices, size_t(d.num_tris) * 3, st);
        normals_.upload(d.normals, d.normals ? size_t(d.num_vertices) * 3 : 0, st);
        uvs_.upload(d.uvs, d.uvs ? size_t(d.num_vertices) * 2 : 0, st);
        tangents_.upload(d.tangents, d.tangents ? size_t(d.num_vertices) * 4 : 0, st);
        textures_.upload(d.textures, size_t(d.num_textures), st);
        texels_.upload(d.texels, size_t(d.num_texels), st);
        srgb_lut_.upload(d.srgb_lut, 256, st);
        punctual_.upload(d.punctual, size_t(d.num_punctual), st);
        upload_lights(d, st);
        phd::DScene& v = view_;
        v.nodes = or_null(nodes_);
        v.indices = or_null(indices_);
        v.normals = or_null(normals_);
        v.uvs = or_null(uvs_);
        v.tangents = or_null(tangents_);
        v.textures = or_null(textures_);
        v.texels = or_null(texels_);
        v.srgb_lut = or_null(srgb_lut_);
        v.punctual = or_null(punctual_);
        v.num_tris = d.num_tris;
        v.num_nodes = d.num_nodes;
        v.top_nodes = top;
        v.num_punctual = d.num_punctual;
        v.num_textures = d.num_textures;
        float dx = d.bounds_max[0] - d.bounds_min[0], dy = d.bounds_max[1] - d.bounds_min[1], dz = d.bounds_max[2] - d.bounds_min[2];
        extent_ = std::fmax(std::sqrt(dx * dx + dy * dy + dz * dz), 1e-3f);
        PH_CHECK(hipStreamSynchronize(st));
        ready_ = true;
    }

    // Re-uploads what a light rebuild changes (GpuRenderer::update_lights, gpu_renderer.cu:299-334).
    void upload_lights(const ph_scene_desc& d, hipStream_t st) {
        tris_.upload(d.tris, size_t(d.num_tris) * 3, st);
        materials_.upload(d.materials, size_t(d.num_materials) * 5, st);
        emissive_.upload(d.emissive, size_t(d.num_emissive), st);
        tri_emissive_.upload(d.tri_emissive, d.tri_emissive ? size_t(d.num_tris) : 0, st);
        alias_.upload(d.emissive_alias, size_t(d.num_emissive), st);
        const size_t env_n = d.env_texels ? size_t(d.env_width) * size_t(d.env_height) : 0;
        env_texels_.upload(d.env_texels, env_n * 3, st);
        env_marginal_.upload(d.env_marginal, env_n ? size_t(d.env_height) + 1 : 0, st);
        env_conditional_.upload(d.env_conditional, env_n ? size_t(d.env_height) * size_t(d.env_width + 1) : 0, st);
        phd::DScene& v = view_;
        v.tris = or_null(tris_);
        v.materials = or_null(materials_);
        v.emissive = or_null(emissive_);
        v.tri_emissive = or_null(tri_emissive_);
        v.alias = or_null(alias_);
        v.num_emissive = d.num_emissive;
        v.emissive_power = d.emissive_power;
        v.env_texels = or_null(env_texels_);
        v.env_marginal = or_null(env_marginal_);
        v.env_conditional = or_null(env_conditional_);
        v.env_w = d.env_width;
        v.env_h = d.env_height;
        // Guide tables for the sampling searches, one entry per power-of-two step of u.
        v.env_guide_marginal = v.env_guide_conditional = nullptr;
        v.env_guide_mh = v.env_guide_mw = 0;
        if (env_n) {
            auto pow2 = [](int n) {
                int m = 1;
                while (m < n && m < (1 << 16)) m <<= 1;
                return m;
            };
            const int mh = pow2(d.env_height), mw = pow2(d.env_width);
            guide_marginal_.allocate(size_t(mh) + 1);
            guide_conditional_.allocate(size_t(d.env_height) * size_t(mw + 1));
            build_cdf_guides(st, env_marginal_.ptr, 1, d.env_height, 0, mh, guide_marginal_.ptr);
            build_cdf_guides(st, env_conditional_.ptr, d.env_height, d.env_width, size_t(d.env_width + 1), mw, guide_conditional_.ptr);
            v.env_guide_marginal = guide_marginal_.ptr;
            v.env_guide_conditional = guide_conditional_.ptr;
            v.env_guide_mh = mh;
            v.env_guide_mw = mw;
        }
        v.env_intensity = d.env_intensity;
        v.env_rotation = d.env_rotation;
        v.env_integral = d.env_marginal_integral;
        std::memcpy(v.env_color, d.env_color, 12);
        v.p_env = d.p_env;
        v.p_emissive = d.p_emissive;
        v.p_punctual = d.p_punctual;
        v.num_materials = d.num_materials;
    }

private:
    DevBuf<float4> nodes_, tris_, materials_;
    DevBuf<int> indices_, tri_emissive_;
    DevBuf<float> normals_, uvs_, tangents_, srgb_lut_, env_texels_, env_marginal_, env_conditional_;
    DevBuf<int> guide_marginal_, guide_conditional_;
    DevBuf<ph_texture_info> textures_;
    DevBuf<uint32_t> texels_;
    DevBuf<ph_emissive_tri> emissive_;
    DevBuf<ph_alias_entry> alias_;
    DevBuf<ph_punctual_light> punctual_;
    phd::DScene view_{};
    float extent_ = 1;
    bool ready_ = false;
};

}  // namespace phk
