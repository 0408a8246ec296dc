// Drives the C++ mirror of the reference's BLAS / TLAS classes (include/vkhrt.hpp).
//   host_api_test nodevice
//       error behaviour without a GPU: Context -> NO_DEVICE, BLASType / geometry pairing
//       (eVoxels with VKHRT_GEOMETRY_VOXELS),
//       CameraLookAt; prints "ok".
//   host_api_test trace <curves.bin> <n> <rays.bin> <m> <hits.bin> [instances]
//   host_api_test batch <curves.bin> <n> <cams.bin> <nviews> <W1xH1,...> <bounces> <seed> <out.bin>
//   host_api_test assetinfo <file.gltf|.glb>
//       ModelLoader::LoadFromFile without a device: prints the scene graph (meshes, nodes with
//       world matrices as hex floats, instances) or the loader's error.
//   host_api_test asset <file> <hair_mode> <rays.bin> <m> <hits.bin>
//       Model::BuildBLASes + TLAS + closest-hit trace of m rays; writes m vkhrt_hit records.
//       builds `instances` phantom BLASes (per-point radius, host pointers) with transforms
//       x -> x + 10 * i, one TLAS, traces m rays (closest hit) and writes m vkhrt_hit records.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "vkhrt.hpp"

static std::vector<float> read_floats(const char* path, size_t n) {
    std::vector<float> v(n);
    FILE* f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot read %zu floats from %s\n", n, path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

static int expect_error(int code, void (*fn)()) {
    try {
        fn();
    } catch (const vkhrt::Error& e) {
        if (e.Code() == code) return 0;
        std::fprintf(stderr, "wrong code %d (%s)\n", e.Code(), e.what());
        return 1;
    }
    std::fprintf(stderr, "no error, expected %d\n", code);
    return 1;
}

static int no_device() {
    int bad = 0;
    bad += expect_error(VKHRT_ERROR_NO_DEVICE, [] { vkhrt::Context ctx(0); });
    bad += expect_error(VKHRT_ERROR_INVALID_ARGUMENT, [] {
        vkhrt::BLASInput in;
        in.type = vkhrt::BLASType::eVoxels;  // voxels need VKHRT_GEOMETRY_VOXELS
        vkhrt::BottomLevelAccelerationStructure b(in, nullptr);
    });
    bad += expect_error(VKHRT_ERROR_INVALID_ARGUMENT, [] {
        vkhrt::BLASInput in;
        in.type = vkhrt::BLASType::eVoxels;
        in.geometry = VKHRT_GEOMETRY_VOXELS;  // paired, but no context
        vkhrt::BottomLevelAccelerationStructure b(in, nullptr);
    });
    bad += expect_error(VKHRT_ERROR_INVALID_ARGUMENT, [] {
        vkhrt::BLASInput in;
        in.type = vkhrt::BLASType::eMesh;  // phantom curves belong to eHair
        vkhrt::BottomLevelAccelerationStructure b(in, nullptr);
    });
    bad += expect_error(VKHRT_ERROR_INVALID_ARGUMENT, [] {
        vkhrt::BLASInput in;  // eHair + phantom, but no context
        vkhrt::BottomLevelAccelerationStructure b(in, nullptr);
    });
    vkhrt_camera cam = vkhrt::CameraLookAt({0, 60, 300}, {0, 40, 0}, 0.57735026f, 16.0f / 9.0f);
    for (int i = 0; i < 32; ++i) std::printf("%a ", double((&cam.view_inverse[0])[i]));
    std::printf("\n%s\n", bad ? "FAILED" : "ok");
    return bad ? 1 : 0;
}

static int trace(int argc, char** argv) {
    if (argc < 7) return 2;
    const size_t n = std::strtoul(argv[3], nullptr, 10), m = std::strtoul(argv[5], nullptr, 10);
    const int ninst = argc > 7 ? std::atoi(argv[7]) : 1;
    std::vector<float> curves = read_floats(argv[2], n * 16), rays = read_floats(argv[4], m * 8);
    auto ctx = std::make_shared<vkhrt::Context>(0);
    std::vector<vkhrt::BottomLevelAccelerationStructure> blases;
    for (int i = 0; i < ninst; ++i) {
        vkhrt::BLASInput in;
        in.type = vkhrt::BLASType::eHair;
        in.geometry = VKHRT_GEOMETRY_PHANTOM_CURVES;
        in.primitives = curves.data();
        in.count = uint32_t(n);
        in.perPointRadius = true;
        in.hostPointers = true;
        in.transform[12] = 10.0f * float(i);  // column-major translation x
        blases.emplace_back(in, ctx);
    }
    vkhrt::TopLevelAccelerationStructure tlas(blases, ctx);
    void *drays = nullptr, *dhits = nullptr;
    if (hipMalloc(&drays, m * sizeof(vkhrt_ray)) != hipSuccess || hipMalloc(&dhits, m * sizeof(vkhrt_hit)) != hipSuccess)
        return 3;
    if (hipMemcpy(drays, rays.data(), m * sizeof(vkhrt_ray), hipMemcpyHostToDevice) != hipSuccess) return 3;
    tlas.TraceRays(static_cast<const vkhrt_ray*>(drays), uint32_t(m), static_cast<vkhrt_hit*>(dhits));
    ctx->Synchronize();
    std::vector<vkhrt_hit> hits(m);
    if (hipMemcpy(hits.data(), dhits, m * sizeof(vkhrt_hit), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    if (hipFree(drays) != hipSuccess || hipFree(dhits) != hipSuccess) return 3;
    FILE* f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(hits.data(), sizeof(vkhrt_hit), m, f) != m) return 4;
    std::fclose(f);
    std::printf("traced %zu rays against %d instance(s) of %zu curves\n", m, ninst, n);
    return 0;
}

// Batched path tracing through the C++ mirror: views from cams.bin (nviews x 32 floats: view
// inverse, projection inverse) with sizes "W1xH1,W2xH2,...", RGBA32F images concatenated to out.
static int batch(int argc, char** argv) {
    if (argc < 9) return 2;
    const size_t n = std::strtoul(argv[3], nullptr, 10), nviews = std::strtoul(argv[5], nullptr, 10);
    std::vector<float> curves = read_floats(argv[2], n * 16), cams = read_floats(argv[4], nviews * 32);
    std::vector<std::pair<uint32_t, uint32_t>> sizes;
    for (const char* p = argv[6]; *p;) {
        char* e = nullptr;
        const uint32_t w = uint32_t(std::strtoul(p, &e, 10));
        const uint32_t h = uint32_t(std::strtoul(e + 1, &e, 10));
        sizes.emplace_back(w, h);
        p = *e ? e + 1 : e;
    }
    if (sizes.size() != nviews) return 2;
    auto ctx = std::make_shared<vkhrt::Context>(0);
    std::vector<vkhrt::BottomLevelAccelerationStructure> blases;
    vkhrt::BLASInput in;
    in.type = vkhrt::BLASType::eHair;
    in.geometry = VKHRT_GEOMETRY_PHANTOM_CURVES;
    in.primitives = curves.data();
    in.count = uint32_t(n);
    in.perPointRadius = true;
    in.hostPointers = true;
    blases.emplace_back(in, ctx);
    vkhrt::TopLevelAccelerationStructure tlas(blases, ctx);
    std::vector<vkhrt_view> views(nviews);
    size_t total = 0;
    for (size_t v = 0; v < nviews; ++v) total += size_t(sizes[v].first) * sizes[v].second;
    void* dimg = nullptr;
    if (hipMalloc(&dimg, total * 16) != hipSuccess) return 3;
    size_t off = 0;
    for (size_t v = 0; v < nviews; ++v) {
        std::memcpy(views[v].camera.view_inverse, &cams[v * 32], 64);
        std::memcpy(views[v].camera.proj_inverse, &cams[v * 32 + 16], 64);
        views[v].width = sizes[v].first;
        views[v].height = sizes[v].second;
        views[v].image = static_cast<char*>(dimg) + off * 16;
        views[v].primitive_ids = nullptr;
        off += size_t(sizes[v].first) * sizes[v].second;
    }
    vkhrt_batch_desc d{};
    d.flags = VKHRT_BATCH_FP64_POLISH | VKHRT_BATCH_GRAPH;
    d.bounces = uint32_t(std::atoi(argv[7]));
    d.seed = std::strtoull(argv[8], nullptr, 10);
    d.environment[0] = 0.2f;
    d.environment[1] = 0.3f;
    d.environment[2] = 0.4f;
    tlas.RenderBatch(views, d);
    ctx->Synchronize();
    const vkhrt_batch_stats st = tlas.BatchStats();
    std::vector<float> img(total * 4);
    if (hipMemcpy(img.data(), dimg, total * 16, hipMemcpyDeviceToHost) != hipSuccess || hipFree(dimg) != hipSuccess)
        return 3;
    FILE* f = std::fopen(argv[9], "wb");
    if (!f || std::fwrite(img.data(), 4, img.size(), f) != img.size()) return 4;
    std::fclose(f);
    std::printf("batch: %llu paths, %llu bounce rays, %llu shadow rays\n", (unsigned long long)st.paths,
                (unsigned long long)st.bounce_rays, (unsigned long long)st.shadow_rays);
    return 0;
}

static int asset_info(const char* path) {
    std::shared_ptr<vkhrt::Model> model;
    try {
        model = vkhrt::ModelLoader().LoadFromFile(path);
    } catch (const vkhrt::Error& e) {
        std::printf("error %d %s\n", e.Code(), e.what());
        return 0;
    }
    const vkhrt::SceneGraph& g = *model->sceneGraph;
    for (const vkhrt::Mesh& m : g.meshes) {
        unsigned long long sum = 0;  // order-sensitive digest of the index array
        for (uint32_t i = 0; i < m.indexCount; ++i) sum = sum * 1000003ull + m.indices[i];
        std::printf("mesh %d %u %u %d %llu %a %a\n", m.primitiveType == vkhrt::Mesh::PrimitiveType::eLines ? 1 : 0,
                    m.vertexCount, m.indexCount, m.material, sum, double(m.boundingBox.min[0]),
                    double(m.boundingBox.max[2]));
    }
    for (const vkhrt::Node& n : g.nodes) {
        std::printf("node %d %zu", n.parent, n.meshes.size());
        const std::array<float, 16> w = n.GetWorldMatrix();
        for (float x : w) std::printf(" %a", double(x));
        std::printf("\n");
    }
    std::printf("materials %zu\nok\n", g.materials.size());
    return 0;
}

static int asset_trace(int argc, char** argv) {
    if (argc < 7) return 2;
    const int mode = std::atoi(argv[3]);
    const size_t m = std::strtoul(argv[5], nullptr, 10);
    std::vector<float> rays = read_floats(argv[4], m * 8);
    auto ctx = std::make_shared<vkhrt::Context>(0);
    auto model = vkhrt::ModelLoader().LoadFromFile(argv[2]);
    std::vector<vkhrt::BottomLevelAccelerationStructure> blases =
        model->BuildBLASes(ctx, vkhrt_geometry_mode(mode), 0.02f, 1.0f, 0);
    vkhrt::TopLevelAccelerationStructure tlas(blases, ctx);
    void *drays = nullptr, *dhits = nullptr;
    if (hipMalloc(&drays, m * sizeof(vkhrt_ray)) != hipSuccess || hipMalloc(&dhits, m * sizeof(vkhrt_hit)) != hipSuccess)
        return 3;
    if (hipMemcpy(drays, rays.data(), m * sizeof(vkhrt_ray), hipMemcpyHostToDevice) != hipSuccess) return 3;
    tlas.TraceRays(static_cast<const vkhrt_ray*>(drays), uint32_t(m), static_cast<vkhrt_hit*>(dhits));
    ctx->Synchronize();
    std::vector<vkhrt_hit> hits(m);
    if (hipMemcpy(hits.data(), dhits, m * sizeof(vkhrt_hit), hipMemcpyDeviceToHost) != hipSuccess) return 3;
    if (hipFree(drays) != hipSuccess || hipFree(dhits) != hipSuccess) return 3;
    FILE* f = std::fopen(argv[6], "wb");
    if (!f || std::fwrite(hits.data(), sizeof(vkhrt_hit), m, f) != m) return 4;
    std::fclose(f);
    std::printf("traced %zu rays against %zu BLASes of %s\n", m, blases.size(), argv[2]);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "nodevice") == 0) return no_device();
    if (argc >= 2 && std::strcmp(argv[1], "trace") == 0) return trace(argc, argv);
    if (argc >= 2 && std::strcmp(argv[1], "batch") == 0) return batch(argc, argv);
    if (argc >= 3 && std::strcmp(argv[1], "assetinfo") == 0) return asset_info(argv[2]);
    if (argc >= 2 && std::strcmp(argv[1], "asset") == 0) return asset_trace(argc, argv);
    std::fprintf(stderr,
                 "usage: host_api_test nodevice | trace curves n rays m hits [instances]\n"
                 "       | batch curves n cams nviews WxH,... bounces seed out\n");
    return 2;
}
