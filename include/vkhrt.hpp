/*
 * C++ host interface over the vkhrt C-ABI, mirroring the reference's classes:
 *   BLASType, BLASInput, BottomLevelAccelerationStructure
 *       (include/bottom_level_acceleration_structure.hpp:12-50)
 *   TopLevelAccelerationStructure (include/top_level_acceleration_structure.hpp:9-23)
 *   ModelLoader, Model, SceneGraph, Node, Mesh (include/resources/model/model_loader.hpp:12-30,
 *       include/resources/model/model.hpp:6-111) over the vkhrt_asset_* loader
 *   VulkanContext -> vkhrt::Context (one device, one stream per call)
 * The Vulkan geometry / build-range structures of BLASInput become a primitive pointer and a
 * count; the type/geometry pairing follows Renderer::InitializeBLAS
 * (source/renderer.cpp:543-692): eHair = phantom curves (AABBs + hair_intersection.rint),
 * eMesh = LSS capsules or DOTS triangles, eVoxels is not on this path.
 * Errors: the reference aborts through VkCheckResult (source/vk_common.cpp:5-15); here a failed
 * call throws vkhrt::Error carrying the vkhrt_status code and vkhrt_last_error().
 * Header-only; link with libvkhrt.so. No HIP headers are needed: streams are void*.
 */
#ifndef VKHRT_HPP
#define VKHRT_HPP

#include <algorithm>
#include <array>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vkhrt.h"

namespace vkhrt {

inline const char* StatusName(int code) {
    switch (code) {
        case VKHRT_SUCCESS: return "SUCCESS";
        case VKHRT_ERROR_INVALID_ARGUMENT: return "INVALID_ARGUMENT";
        case VKHRT_ERROR_OUT_OF_MEMORY: return "OUT_OF_MEMORY";
        case VKHRT_ERROR_DEVICE: return "DEVICE";
        case VKHRT_ERROR_UNSUPPORTED: return "UNSUPPORTED";
        case VKHRT_ERROR_STACK_OVERFLOW: return "STACK_OVERFLOW";
        case VKHRT_ERROR_NO_DEVICE: return "NO_DEVICE";
        default: return "UNKNOWN";
    }
}

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what)
        : std::runtime_error(what + " [" + StatusName(code) + "]"), _code(code) {}
    int Code() const noexcept { return _code; }

private:
    int _code;
};

class Context {
public:
    explicit Context(int device = 0) {
        int rc = vkhrt_context_create(device, &_handle);
        if (rc != VKHRT_SUCCESS) throw Error(rc, "vkhrt_context_create");
    }
    ~Context() { vkhrt_context_destroy(_handle); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    vkhrt_context* Handle() const noexcept { return _handle; }
    void Check(int rc, const char* what) const {
        if (rc != VKHRT_SUCCESS) throw Error(rc, std::string(what) + ": " + vkhrt_last_error(_handle));
    }
    void SetOption(int option, int value) { Check(vkhrt_set_option(_handle, option, value), "vkhrt_set_option"); }
    void Synchronize(void* stream = nullptr) { Check(vkhrt_synchronize(_handle, stream), "vkhrt_synchronize"); }

private:
    vkhrt_context* _handle = nullptr;
};

// include/bottom_level_acceleration_structure.hpp:12-17
enum class BLASType : uint8_t {
    eMesh,
    eHair,
    eVoxels,
};

// include/bottom_level_acceleration_structure.hpp:19-29. `transform` is a column-major 4x4
// (glm::mat4); `primitives` holds `count` records in the layout of vkhrt_blas_desc.
struct BLASInput {
    BLASType type = BLASType::eHair;
    std::array<float, 16> transform{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    vkhrt_geometry_mode geometry = VKHRT_GEOMETRY_PHANTOM_CURVES;
    const float* primitives = nullptr;
    uint32_t count = 0;
    float radius = 0.02f;  // hair_intersection.rint:18
    bool perPointRadius = false;
    bool hostPointers = false;
    uint32_t materialIndex = 0;
    uint32_t leafSize = 0;
    const uint32_t* indices = nullptr;  // VKHRT_GEOMETRY_TRIANGLES: 3 per triangle (count = triangles)
    uint32_t vertexCount = 0;           // VKHRT_GEOMETRY_TRIANGLES: vertices in `primitives`
    float voxelSize = 0.1f;             // VKHRT_GEOMETRY_VOXELS (eVoxels): cell edge
    uint32_t voxelLod = 0;              // VKHRT_GEOMETRY_VOXELS: level of detail traced
    bool voxelReferenceSweep = false;   // VKHRT_GEOMETRY_VOXELS: the reference's fill loop
    bool hasBounds = false;             // VKHRT_GEOMETRY_VOXELS: grid over boundsMin..boundsMax
    std::array<float, 3> boundsMin{}, boundsMax{};
};

class BottomLevelAccelerationStructure {
public:
    BottomLevelAccelerationStructure(const BLASInput& input, const std::shared_ptr<Context>& context,
                                     void* stream = nullptr)
        : _type(input.type), _transform(input.transform), _context(context) {
        const bool hair = input.geometry == VKHRT_GEOMETRY_PHANTOM_CURVES;
        const bool voxels = input.geometry == VKHRT_GEOMETRY_VOXELS;
        if ((input.type == BLASType::eHair) != hair || (input.type == BLASType::eVoxels) != voxels)
            throw Error(VKHRT_ERROR_INVALID_ARGUMENT,
                        "BottomLevelAccelerationStructure: eHair pairs with phantom curves, eVoxels with voxels, "
                        "eMesh with LSS / DOTS / triangles");
        if (!context) throw Error(VKHRT_ERROR_INVALID_ARGUMENT, "BottomLevelAccelerationStructure: no context");
        vkhrt_blas_desc d{};
        d.mode = uint32_t(input.geometry);
        d.flags = VKHRT_BLAS_HAS_TRANSFORM | (input.perPointRadius ? VKHRT_BLAS_PER_POINT_RADIUS : 0u) |
                  (input.hostPointers ? VKHRT_BLAS_HOST_POINTERS : 0u);
        d.primitives = input.primitives;
        d.count = input.count;
        d.radius = input.radius;
        d.material_index = input.materialIndex;
        d.leaf_size = input.leafSize;
        d.indices = input.indices;
        d.vertex_count = input.vertexCount;
        d.voxel_size = input.voxelSize;
        d.voxel_lod = input.voxelLod;
        d.flags |= (input.voxelReferenceSweep ? VKHRT_BLAS_VOXEL_REFERENCE_SWEEP : 0u) |
                   (input.hasBounds ? VKHRT_BLAS_HAS_BOUNDS : 0u);
        for (int k = 0; k < 3; ++k) {
            d.bounds_min[k] = input.boundsMin[k];
            d.bounds_max[k] = input.boundsMax[k];
        }
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) d.transform[r * 4 + c] = input.transform[c * 4 + r];
        _context->Check(vkhrt_blas_build(_context->Handle(), &d, stream, &_blas), "vkhrt_blas_build");
    }
    // Takes ownership of a BLAS made by another entry point (e.g. vkhrt_asset_build).
    static BottomLevelAccelerationStructure Adopt(vkhrt_blas* blas, BLASType type,
                                                  const std::array<float, 16>& transform,
                                                  const std::shared_ptr<Context>& context) {
        return BottomLevelAccelerationStructure(blas, type, transform, context);
    }
    ~BottomLevelAccelerationStructure() { vkhrt_blas_destroy(_blas); }
    BottomLevelAccelerationStructure(BottomLevelAccelerationStructure&& other) noexcept
        : _blas(std::exchange(other._blas, nullptr)),
          _type(other._type),
          _transform(other._transform),
          _context(std::move(other._context)) {}
    BottomLevelAccelerationStructure& operator=(BottomLevelAccelerationStructure&&) = delete;
    BottomLevelAccelerationStructure(const BottomLevelAccelerationStructure&) = delete;
    BottomLevelAccelerationStructure& operator=(const BottomLevelAccelerationStructure&) = delete;

    const vkhrt_blas* Structure() const noexcept { return _blas; }
    BLASType Type() const noexcept { return _type; }
    const std::array<float, 16>& Transform() const noexcept { return _transform; }
    vkhrt_blas_info Info() const {
        vkhrt_blas_info info{};
        _context->Check(vkhrt_blas_info_get(_blas, &info), "vkhrt_blas_info_get");
        return info;
    }

private:
    BottomLevelAccelerationStructure(vkhrt_blas* blas, BLASType type, const std::array<float, 16>& transform,
                                     const std::shared_ptr<Context>& context)
        : _blas(blas), _type(type), _transform(transform), _context(context) {}

    vkhrt_blas* _blas = nullptr;
    BLASType _type;
    std::array<float, 16> _transform;
    std::shared_ptr<Context> _context;
};

class TopLevelAccelerationStructure {
public:
    // Instance i = blases[i] (customIndex = i, mask 0xFF), source/top_level_acceleration_structure.cpp:23-43.
    TopLevelAccelerationStructure(const std::vector<BottomLevelAccelerationStructure>& blases,
                                  const std::shared_ptr<Context>& context, void* stream = nullptr)
        : _context(context) {
        if (!context) throw Error(VKHRT_ERROR_INVALID_ARGUMENT, "TopLevelAccelerationStructure: no context");
        std::vector<const vkhrt_blas*> handles;
        handles.reserve(blases.size());
        for (const auto& b : blases) handles.push_back(b.Structure());
        _context->Check(vkhrt_tlas_build(_context->Handle(), handles.data(), uint32_t(handles.size()), stream, &_tlas),
                        "vkhrt_tlas_build");
    }
    ~TopLevelAccelerationStructure() { vkhrt_tlas_destroy(_tlas); }
    TopLevelAccelerationStructure(const TopLevelAccelerationStructure&) = delete;
    TopLevelAccelerationStructure& operator=(const TopLevelAccelerationStructure&) = delete;
    TopLevelAccelerationStructure(TopLevelAccelerationStructure&&) = delete;
    TopLevelAccelerationStructure& operator=(TopLevelAccelerationStructure&&) = delete;

    const vkhrt_tlas* Structure() const noexcept { return _tlas; }

    // traceRayEXT over `count` device ray records (shaders/ray_gen.rgen:26-45).
    void TraceRays(const vkhrt_ray* rays, uint32_t count, vkhrt_hit* hits, uint32_t flags = 0,
                   void* stream = nullptr) const {
        _context->Check(vkhrt_trace_rays(_context->Handle(), _tlas, rays, count, hits, flags, stream),
                        "vkhrt_trace_rays");
    }
    // traceRaysKHR(W, H, 1) with the closest-hit shading (source/renderer.cpp:156-166).
    void Render(const vkhrt_camera& camera, const vkhrt_render_desc& desc, void* stream = nullptr) const {
        _context->Check(vkhrt_render(_context->Handle(), _tlas, &camera, &desc, stream), "vkhrt_render");
    }
    // Extensions (no reference counterpart): batched path tracing of many views, its stats, the
    // autotuner on a render or batch workload, and MLP training rays from the hair surface.
    void RenderBatch(const std::vector<vkhrt_view>& views, const vkhrt_batch_desc& desc, void* stream = nullptr) const {
        _context->Check(vkhrt_render_batch(_context->Handle(), _tlas, views.data(), uint32_t(views.size()), &desc, stream),
                        "vkhrt_render_batch");
    }
    vkhrt_batch_stats BatchStats() const {
        vkhrt_batch_stats st{};
        _context->Check(vkhrt_batch_stats_get(_context->Handle(), &st), "vkhrt_batch_stats_get");
        return st;
    }
    vkhrt_tuning AutotuneRender(const vkhrt_camera& camera, const vkhrt_render_desc& desc, uint32_t repeats = 3,
                                void* stream = nullptr) const {
        vkhrt_tuning t{};
        _context->Check(vkhrt_autotune_render(_context->Handle(), _tlas, &camera, &desc, repeats, &t, stream),
                        "vkhrt_autotune_render");
        return t;
    }
    vkhrt_tuning AutotuneBatch(const std::vector<vkhrt_view>& views, const vkhrt_batch_desc& desc,
                               uint32_t repeats = 3, void* stream = nullptr) const {
        vkhrt_tuning t{};
        _context->Check(vkhrt_autotune_batch(_context->Handle(), _tlas, views.data(), uint32_t(views.size()), &desc,
                                             repeats, &t, stream),
                        "vkhrt_autotune_batch");
        return t;
    }
    void SurfaceShadowRays(const vkhrt_area_light* light, uint32_t count, uint64_t seed, vkhrt_ray* rays,
                           void* stream = nullptr) const {
        _context->Check(vkhrt_generate_surface_shadow_rays(_context->Handle(), _tlas, light, count, seed, rays, stream),
                        "vkhrt_generate_surface_shadow_rays");
    }

private:
    vkhrt_tlas* _tlas = nullptr;
    std::shared_ptr<Context> _context;
};

// ---- models: include/resources/model/model.hpp (Node :6-16, AABB :18-22, Mesh :42-66,
// SceneGraph :98-111) and ModelLoader (model_loader.hpp:12-30). Each mesh keeps its own
// vertex and index arrays (mesh-local indices), a node names its parent by index, and the
// loader throws Error where the reference logs and returns nullptr (model_loader.cpp:280-284).
struct AABB {
    std::array<float, 3> min{};
    std::array<float, 3> max{};
};

struct Node {
    std::string name;
    int32_t parent = -1;               // index into SceneGraph::nodes, -1 for the root
    std::array<float, 16> localMatrix{};  // column-major (glm::mat4)
    std::array<float, 16> worldMatrix{};  // parents applied from the nearest (model.cpp:6-18)
    std::vector<uint32_t> meshes;
    std::array<float, 16> GetWorldMatrix() const { return worldMatrix; }
};

struct Mesh {
    enum class PrimitiveType : uint8_t {
        eTriangles,
        eLines,
    };
    PrimitiveType primitiveType = PrimitiveType::eTriangles;
    uint32_t vertexCount = 0;
    uint32_t indexCount = 0;
    const float* vertices = nullptr;    // vertexCount x Vertex (position, normal, texCoord), host
    const uint32_t* indices = nullptr;  // mesh-local, host
    int32_t material = -1;              // index into SceneGraph::materials, -1 = none
    AABB boundingBox;
    std::string name;
    uint32_t GetIndicesPerFaceNum() const { return primitiveType == PrimitiveType::eLines ? 2u : 3u; }
};

struct SceneGraph {
    std::vector<Node> nodes;  // pre-order, nodes[0] is the root
    std::vector<Mesh> meshes;
    std::vector<vkhrt_asset_material> materials;
};

class Model {
public:
    explicit Model(vkhrt_asset* asset) : _asset(asset), sceneGraph(std::make_shared<SceneGraph>()) {
        vkhrt_asset_info info{};
        vkhrt_asset_info_get(_asset, &info);
        for (uint32_t i = 0; i < info.mesh_count; ++i) {
            vkhrt_asset_mesh m{};
            vkhrt_asset_mesh_get(_asset, i, &m);
            Mesh& M = sceneGraph->meshes.emplace_back();
            M.primitiveType = m.primitive == VKHRT_PRIMITIVE_LINES ? Mesh::PrimitiveType::eLines
                                                                   : Mesh::PrimitiveType::eTriangles;
            M.vertexCount = m.vertex_count;
            M.indexCount = m.index_count;
            M.vertices = m.vertices;
            M.indices = m.indices;
            M.material = m.material;
            for (int k = 0; k < 3; ++k) {
                M.boundingBox.min[k] = m.bounds_min[k];
                M.boundingBox.max[k] = m.bounds_max[k];
            }
            M.name = m.name ? m.name : "";
        }
        for (uint32_t i = 0; i < info.node_count; ++i) {
            vkhrt_asset_node n{};
            vkhrt_asset_node_get(_asset, i, &n);
            Node& N = sceneGraph->nodes.emplace_back();
            N.name = n.name ? n.name : "";
            N.parent = n.parent;
            std::copy(n.local, n.local + 16, N.localMatrix.begin());
            std::copy(n.world, n.world + 16, N.worldMatrix.begin());
            N.meshes.assign(n.meshes, n.meshes + n.mesh_count);
        }
        for (uint32_t i = 0; i < info.material_count; ++i) {
            vkhrt_asset_material m{};
            vkhrt_asset_material_get(_asset, i, &m);
            sceneGraph->materials.push_back(m);
        }
    }
    ~Model() { vkhrt_asset_destroy(_asset); }
    Model(const Model&) = delete;
    Model& operator=(const Model&) = delete;

    const vkhrt_asset* Handle() const noexcept { return _asset; }

    // ProcessHairCurves / ProcessHairLSS / ProcessHairDOTS on the device and
    // Renderer::InitializeBLAS for this model (renderer.cpp:694-728): one BLAS per (node, mesh)
    // in node pre-order, line meshes as `hairMode`, triangle meshes as eMesh triangles.
    std::vector<BottomLevelAccelerationStructure> BuildBLASes(const std::shared_ptr<Context>& context,
                                                              vkhrt_geometry_mode hairMode = VKHRT_GEOMETRY_PHANTOM_CURVES,
                                                              float radius = 0.02f, float tension = 1.0f,
                                                              uint32_t defaultMaterial = 0, void* stream = nullptr,
                                                              float voxelSize = 0.1f, uint32_t voxelLod = 0) const {
        if (!context) throw Error(VKHRT_ERROR_INVALID_ARGUMENT, "Model::BuildBLASes: no context");
        vkhrt_asset_build_desc d{};
        d.hair_mode = uint32_t(hairMode);
        d.radius = radius;
        d.tension = tension;
        d.default_material = defaultMaterial;
        d.voxel_size = voxelSize;
        d.voxel_lod = voxelLod;
        uint32_t n = 0;
        context->Check(vkhrt_asset_build(context->Handle(), _asset, &d, stream, nullptr, 0, &n), "vkhrt_asset_build");
        std::vector<vkhrt_blas*> raw(n, nullptr);
        context->Check(vkhrt_asset_build(context->Handle(), _asset, &d, stream, raw.data(), n, &n), "vkhrt_asset_build");
        std::vector<BottomLevelAccelerationStructure> out;
        out.reserve(n);
        for (uint32_t i = 0; i < n; ++i) {
            vkhrt_asset_instance inst{};
            vkhrt_asset_instance_get(_asset, i, &inst);
            const Mesh& M = sceneGraph->meshes[inst.mesh];
            const bool lines = M.primitiveType == Mesh::PrimitiveType::eLines;
            const BLASType type = lines && hairMode == VKHRT_GEOMETRY_PHANTOM_CURVES ? BLASType::eHair
                                  : lines && hairMode == VKHRT_GEOMETRY_VOXELS       ? BLASType::eVoxels
                                                                                     : BLASType::eMesh;
            out.push_back(BottomLevelAccelerationStructure::Adopt(raw[i], type, sceneGraph->nodes[inst.node].worldMatrix,
                                                                  context));
        }
        return out;
    }

private:
    vkhrt_asset* _asset = nullptr;

public:
    std::shared_ptr<SceneGraph> sceneGraph;
};

class ModelLoader {
public:
    // .gltf (external buffers or data: URIs) or .glb.
    std::shared_ptr<Model> LoadFromFile(const std::string& path) const {
        vkhrt_asset* a = nullptr;
        int rc = vkhrt_asset_load(path.c_str(), &a);
        if (rc != VKHRT_SUCCESS) throw Error(rc, "ModelLoader::LoadFromFile: " + std::string(vkhrt_asset_last_error()));
        return std::make_shared<Model>(a);
    }
};

// FlyCamera view/projection (source/fly_camera.cpp:25-35), inverted as in
// Renderer::UpdateCameraResource (source/renderer.cpp:189-195).
inline vkhrt_camera CameraLookAt(const std::array<float, 3>& eye, const std::array<float, 3>& center,
                                 float tanHalfFovy, float aspect, const std::array<float, 3>& up = {0, 1, 0},
                                 float znear = 0.1f, float zfar = 1000.0f) {
    vkhrt_camera cam{};
    int rc = vkhrt_camera_look_at(eye.data(), center.data(), up.data(), tanHalfFovy, aspect, znear, zfar, &cam);
    if (rc != VKHRT_SUCCESS) throw Error(rc, "vkhrt_camera_look_at");
    return cam;
}

}  // namespace vkhrt

#endif /* VKHRT_HPP */
