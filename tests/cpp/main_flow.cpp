// The reference driver's flow (main.cu:9-107) written against the
// reference-compatible C++ API (include/xr/reconstruct.hpp): init_params, raw
// sinograms loaded by their reference file names and clamped, XTT::reconstruct
// (which also writes loss.csv), then the coefficient volumes, the main
// direction and its angles saved under the reference's file names.
//
// usage: main_flow settings.json tag epochs batch      run the flow
//        main_flow settings.json tag --paths           print the file names (no GPU)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "xr/reconstruct.hpp"

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s settings.json tag (epochs batch | --paths)\n", argv[0]);
        return 2;
    }
    init_params(argv[2], argv[1]);
    const int C = NUM_PROJ_COND, K = NUM_BASIS_VECTOR;
    if (std::strcmp(argv[3], "--paths") == 0) {
        for (int i = 0; i < C; i++) std::printf("load %s\n", xr::sinogram_load_path(i).c_str());
        for (int i = 0; i < C; i++) std::printf("proj %s\n", xr::sinogram_save_path(i).c_str());
        for (int k = 0; k < K; k++) std::printf("volume %s\n", xr::volume_save_path(k).c_str());
        for (int i = 0; i < 3; i++) std::printf("direction %s\n", xr::direction_save_path(i).c_str());
        for (int i = 0; i < 2; i++) std::printf("angle %s\n", xr::angle_save_path(i).c_str());
        std::printf("loss %s\n", xr_detail::loss_path().c_str());
        return 0;
    }
    if (argc < 5) return 2;
    const int epochs = std::atoi(argv[3]), batch = std::atoi(argv[4]);
    std::vector<Volume<float>> sinogram(C), ct(K), md(3), angle(2);
    for (auto& e : sinogram) e = Volume<float>(NUM_DETECT_U, NUM_DETECT_V, NUM_PROJ);
    for (auto& e : ct) e = Volume<float>(NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
    for (auto& e : md) e = Volume<float>(NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
    for (auto& e : angle) e = Volume<float>(NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
    Geometry geom(SRC_DETECT_DISTANCE, SRC_OBJ_DISTANCE, DETECTOR_SIZE, NUM_VOXEL, NUM_DETECT_U, NUM_PROJ);
    for (int i = 0; i < C; i++) {
        sinogram[i].load(xr::sinogram_load_path(i), NUM_DETECT_U, NUM_DETECT_V, NUM_PROJ);
        sinogram[i].forEach([](float value) -> float { return value < 0.0 ? 1e-8f : value; });  // main.cu:35
    }
    for (auto& e : ct) e.forEach([](float) -> float { return 0.01f; });  // main.cu:40-42
    XTT::reconstruct(sinogram.data(), ct.data(), md.data(), geom, epochs, batch, Rotate::CW, Method::MLEM, 1e-3f);
    for (int i = 0; i < C; i++) sinogram[i].save(xr::sinogram_save_path(i));
    for (int k = 0; k < K; k++) ct[k].save(xr::volume_save_path(k));
    for (int i = 0; i < 3; i++) md[i].save(xr::direction_save_path(i));
    calcAngleFromMD(md.data(), angle.data(), NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
    for (int i = 0; i < 2; i++) angle[i].save(xr::angle_save_path(i));
    directionNormalization(md.data());
    double s = 0;
    for (int i = 0; i < 3; i++)
        for (int64_t j = 0; j < (int64_t)NUM_VOXEL * NUM_VOXEL * NUM_VOXEL; j++) s += md[i].get()[j] * md[i].get()[j];
    std::printf("{\"unit_directions\": %.1f}\n", s);
    return 0;
}
