// Runs the device BVH build (photon_amd/csrc/hip/bvh_build.hip, its host driver and its kernels) on the
// CPU through hip/hip_runtime.h, for tests under the sanitizers.
//
// usage: bvh_build_cpu IN OUT
//   IN:  int32 num_vertices, int32 num_tris, float positions[3 * num_vertices], int32 indices[3 * num_tris]
//   OUT: int64 node count, ph_bvh_node nodes[count], int32 order[num_tris], ph_bvh_build_stats
// The build parameters are the ones ph_bvh_build passes for its defaults (32 bins, leaves of up to 8
// triangles, traversal cost 1).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "photon_hip.h"

namespace phk {
int64_t bvh_build_device(const float* positions, int num_vertices, const int* indices, int n, int bins, int max_leaf, float traversal_cost,
                         ph_bvh_node* nodes_out, int* order_out, ph_bvh_build_stats& stats);
}

static void read_all(FILE* f, void* p, size_t bytes) {
    if (bytes && std::fread(p, 1, bytes, f) != bytes) {
        std::fprintf(stderr, "short input\n");
        std::exit(2);
    }
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s IN OUT\n", argv[0]);
        return 2;
    }
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int nv = 0, nt = 0;
    read_all(in, &nv, 4);
    read_all(in, &nt, 4);
    std::vector<float> pos(size_t(nv) * 3);
    std::vector<int> idx(size_t(nt) * 3);
    read_all(in, pos.data(), pos.size() * 4);
    read_all(in, idx.data(), idx.size() * 4);
    std::fclose(in);
    std::vector<ph_bvh_node> nodes(size_t(nt > 0 ? 2 * nt - 1 : 1));
    std::vector<int> order(static_cast<size_t>(nt));
    ph_bvh_build_stats st{};
    const int64_t count = phk::bvh_build_device(nv ? pos.data() : nullptr, nv, nt ? idx.data() : nullptr, nt, 32, 8, 1.0f, nodes.data(),
                                                nt ? order.data() : nullptr, st);
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(&count, 8, 1, out);
    std::fwrite(nodes.data(), sizeof(ph_bvh_node), size_t(count), out);
    std::fwrite(order.data(), 4, order.size(), out);
    std::fwrite(&st, sizeof(st), 1, out);
    std::fclose(out);
    std::printf("%lld nodes, %.1f ms\n", (long long)count, st.build_ms);
    return 0;
}
