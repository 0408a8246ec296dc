// Loads every file named on the command line with the asset reader (vkhrt_amd/csrc/asset.cpp,
// compiled into this host-only program) and walks the result through the C-ABI getters.
// Built with AddressSanitizer and UndefinedBehaviorSanitizer by tests/cpp/Makefile, so an
// out-of-bounds read or undefined arithmetic on a malformed file stops the program.
#include <cstdio>

#include "vkhrt.h"

int main(int argc, char** argv) {
    int loaded = 0, failed = 0;
    for (int i = 1; i < argc; ++i) {
        vkhrt_asset* a = nullptr;
        if (vkhrt_asset_load(argv[i], &a) != VKHRT_SUCCESS) {
            ++failed;
            continue;
        }
        vkhrt_asset_info info{};
        vkhrt_asset_info_get(a, &info);
        double sink = 0.0;
        for (uint32_t m = 0; m < info.mesh_count; ++m) {
            vkhrt_asset_mesh mesh{};
            vkhrt_asset_mesh_get(a, m, &mesh);
            for (uint32_t k = 0; k < mesh.vertex_count * 8u; ++k) sink += mesh.vertices[k];
            for (uint32_t k = 0; k < mesh.index_count; ++k) sink += mesh.vertices[mesh.indices[k] * 8u];
        }
        for (uint32_t n = 0; n < info.node_count; ++n) {
            vkhrt_asset_node node{};
            vkhrt_asset_node_get(a, n, &node);
            for (uint32_t k = 0; k < node.mesh_count; ++k) sink += node.meshes[k];
        }
        for (uint32_t k = 0; k < info.instance_count; ++k) {
            vkhrt_asset_instance inst{};
            vkhrt_asset_instance_get(a, k, &inst);
            sink += inst.transform[0];
        }
        vkhrt_asset_destroy(a);
        ++loaded;
        if (sink == 1e300) std::printf(" ");
    }
    std::printf("%d loaded, %d rejected\n", loaded, failed);
    return 0;
}
