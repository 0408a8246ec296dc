// In-memory form of a loaded glTF 2.0 asset (asset.cpp): the SceneGraph that
// ModelLoader::LoadModel builds (source/resources/model/model_loader.cpp:293-312) with the
// per-mesh vertex and index arrays kept apart, so that every index is mesh-local.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "vkhrt.h"

namespace vk {

struct AssetMesh {
    uint32_t primitive = VKHRT_PRIMITIVE_TRIANGLES;
    int32_t material = -1;
    std::vector<float> vertices;    // 8 floats per vertex: position, normal, texCoord
    std::vector<uint32_t> indices;  // 2 per line / 3 per triangle, below vertices.size() / 8
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    std::string name;
};

struct AssetNode {
    std::string name;
    int32_t parent = -1;
    std::vector<uint32_t> meshes;
    float local[16];  // column-major
    float world[16];  // column-major, parent chain applied
};

struct AssetInstance {
    uint32_t node, mesh;
};

// Node::GetWorldMatrix (source/resources/model/model.cpp:6-18) over float matrices.
void world_matrix(const std::vector<AssetNode>& nodes, uint32_t i, float out[16]);

}  // namespace vk

struct vkhrt_asset {
    std::vector<vk::AssetMesh> meshes;
    std::vector<vk::AssetNode> nodes;  // pre-order, nodes[0] is the root
    std::vector<vkhrt_asset_material> materials;
    std::vector<vk::AssetInstance> instances;  // (node, mesh) in Renderer::InitializeBLAS order
};
