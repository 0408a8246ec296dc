// Drives the C++ mirror of the reference's BLAS / TLAS classes (include/vkhrt.hpp).
//   host_api_test nodevice
//       error behaviour without a GPU: Context -> NO_DEVICE, BLASType / geometry pairing,
//       CameraLookAt; prints "ok".
//   host_api_test trace <curves.bin> <n> <rays.bin> <m> <hits.bin> [instances]
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
    bad += expect_error(VKHRT_ERROR_UNSUPPORTED, [] {
        vkhrt::BLASInput in;
        in.type = vkhrt::BLASType::eVoxels;
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

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "nodevice") == 0) return no_device();
    if (argc >= 2 && std::strcmp(argv[1], "trace") == 0) return trace(argc, argv);
    std::fprintf(stderr, "usage: host_api_test nodevice | trace curves n rays m hits [instances]\n");
    return 2;
}
