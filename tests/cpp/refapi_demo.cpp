// Drop-in check of the reference-compatible C++ API (include/xr/reconstruct.hpp):
// the same calls the reference's main.cu makes (init_params, Volume, Geometry,
// IR / FDK / XTT ::reconstruct, IR::gradReconstruct), on a small synthetic problem.  Prints one
// JSON line with checksums for the Python test to compare.
#include <cstdio>

#include "xr/reconstruct.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s settings.json tag\n", argv[0]);
        return 2;
    }
    init_params(argv[2], argv[1]);
    Geometry geom(SRC_DETECT_DISTANCE, SRC_OBJ_DISTANCE, DETECTOR_SIZE, NUM_VOXEL, NUM_DETECT_U, NUM_PROJ);
    const int C = NUM_PROJ_COND, K = NUM_BASIS_VECTOR, N = NUM_VOXEL;
    std::vector<Volume<float>> sino(C), ct(K), md(3), vol(C);
    for (int c = 0; c < C; c++) {
        sino[c] = Volume<float>(NUM_DETECT_U, NUM_DETECT_V, NUM_PROJ);
        // smooth positive synthetic data
        for (int n = 0; n < NUM_PROJ; n++)
            for (int v = 0; v < NUM_DETECT_V; v++)
                for (int u = 0; u < NUM_DETECT_U; u++)
                    sino[c](u, v, n) = 0.05f + 0.01f * (float)((u * 7 + v * 3 + n + c) % 11);
        vol[c] = Volume<float>(N, N, N);
        vol[c].forEach([](float) -> float { return 0.01f; });
    }
    for (auto& e : ct) {
        e = Volume<float>(N, N, N);
        e.forEach([](float) -> float { return 0.01f; });
    }
    IR::reconstruct(sino.data(), vol.data(), geom, 2, 4, Rotate::CW, Method::MLEM);
    double s_ir = 0;
    for (int c = 0; c < C; c++)
        for (int64_t i = 0; i < (int64_t)N * N * N; i++) s_ir += vol[c].get()[i];
    XTT::reconstruct(sino.data(), ct.data(), md.data(), geom, 1, 4, Rotate::CW, Method::MLEM);
    double s_xtt = 0, s_md = 0;
    for (int k = 0; k < K; k++)
        for (int64_t i = 0; i < (int64_t)N * N * N; i++) s_xtt += ct[k].get()[i];
    for (int i = 0; i < 3; i++)
        for (int64_t j = 0; j < (int64_t)N * N * N; j++) s_md += std::fabs(md[i].get()[j]);
    std::vector<Volume<float>> fv(C);
    for (int c = 0; c < C; c++) fv[c] = Volume<float>(N, N, N);
    FDK::reconstruct(sino.data(), fv.data(), geom, Rotate::CW);
    double s_fdk = 0;
    for (int c = 0; c < C; c++)
        for (int64_t i = 0; i < (int64_t)N * N * N; i++) s_fdk += fv[c].get()[i];
    // IR::gradReconstruct from sinogram[0] as FDK left it; its gradient volumes
    // are saved under VOLUME_PATH, taken from argv[3] when given
    double s_grad = 0;
    int voxel_after = 0;
    if (argc > 3) {
        VOLUME_PATH = argv[3];
        Geometry g2 = geom;
        std::vector<Volume<float>> gv(1);
        gv[0] = Volume<float>(N, N, N);
        IR::gradReconstruct(sino.data(), gv.data(), g2, 1, 4, Rotate::CW, Method::ART, 1e-2f);
        for (int64_t i = 0; i < (int64_t)N * N * N; i++) s_grad += gv[0].get()[i];
        voxel_after = g2.voxel;
    }
    std::printf("{\"ir\": %.9g, \"xtt\": %.9g, \"md\": %.9g, \"fdk\": %.9g, \"grad\": %.9g, \"voxel_after\": %d}\n",
                s_ir, s_xtt, s_md, s_fdk, s_grad, voxel_after);
    return 0;
}
