// Reference-compatible C++ host API over the C ABI (include/xr.h).
//
// Mirrors the reference's host interface so callers can switch by changing the
// include and linking libxr.so:
//   reconstruct/include/reconstruct.cuh:14-73   Method, Rotate, IR/XTT/FDK::reconstruct
//   reconstruct/include/geometry.h:8-25         Geometry (voxSize = sod*detSize/sdd)
//   reconstruct/include/volume.h:18-165         Volume<T> (x fastest, raw float32 I/O)
//   reconstruct/include/params.h:9-37           init_params and its globals
//   direction/include/pca.cuh:9-12              calcEigenVector, calcAngleFromMD,
//                                               directionNormalization
// Differences, all deliberate:
//   - errors throw std::runtime_error (the reference returns void and does no
//     error checking);
//   - init_params(tag) reads $XR_SETTINGS or ./settings.json instead of a
//     hard-coded absolute path; init_params(tag, path) takes the path;
//   - Volume's copy constructor/assignment allocate before copying (the
//     reference copies into an unallocated buffer, volume.h:34-46);
//   - K (basis vectors) and the number of conditions have no fixed caps;
//   - XR_LOSS_CSV can move (or, empty, turn off) the loss file the reference
//     always writes to ../python/loss.csv.
// Header only; C++17.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <random>
#include <string>
#include <vector>

#include "../xr.h"

enum class Method { XTT, MLEM, ART, FDK };
enum class Rotate { CW, CCW };

class Geometry {
public:
    Geometry(float sdd, float sod, float detSize, int voxel, int detect, int nProj)
        : sdd(sdd), sod(sod), detSize(detSize), voxel(voxel), detect(detect), nProj(nProj) {
        voxSize = sod * detSize / sdd;
    }
    float sdd;
    float sod;
    float voxSize;
    float detSize;
    int voxel;
    int detect;
    int nProj;
};

template <typename T>
class Volume {
public:
    Volume() : sizeX(0), sizeY(0), sizeZ(0) {}
    explicit Volume(int64_t sizeX, int64_t sizeY, int64_t sizeZ) : sizeX(sizeX), sizeY(sizeY), sizeZ(sizeZ) {
        data = std::make_unique<T[]>(sizeX * sizeY * sizeZ);
    }
    explicit Volume(std::string& filename, int64_t sizeX, int64_t sizeY, int64_t sizeZ) {
        load(filename, sizeX, sizeY, sizeZ);
    }
    Volume(const Volume& v) : sizeX(v.sizeX), sizeY(v.sizeY), sizeZ(v.sizeZ) {
        data = std::make_unique<T[]>(sizeX * sizeY * sizeZ);
        if (v.data) std::memcpy(data.get(), v.data.get(), sizeX * sizeY * sizeZ * sizeof(T));
    }
    Volume& operator=(const Volume& v) {
        if (this == &v) return *this;
        sizeX = v.sizeX, sizeY = v.sizeY, sizeZ = v.sizeZ;
        data = std::make_unique<T[]>(sizeX * sizeY * sizeZ);
        if (v.data) std::memcpy(data.get(), v.data.get(), sizeX * sizeY * sizeZ * sizeof(T));
        return *this;
    }
    Volume(Volume&& v) noexcept : sizeX(v.sizeX), sizeY(v.sizeY), sizeZ(v.sizeZ), data(std::move(v.data)) {
        v.sizeX = v.sizeY = v.sizeZ = 0;
    }
    Volume& operator=(Volume&& v) noexcept {
        sizeX = v.sizeX, sizeY = v.sizeY, sizeZ = v.sizeZ;
        v.sizeX = v.sizeY = v.sizeZ = 0;
        data = std::move(v.data);
        return *this;
    }
    ~Volume() = default;

    T& operator()(int64_t x, int64_t y, int64_t z) { return data[z * (sizeX * sizeY) + y * sizeX + x]; }
    T operator()(int64_t x, int64_t y, int64_t z) const { return data[z * (sizeX * sizeY) + y * sizeX + x]; }
    T* get() const { return data.get(); }

    T mean() const {
        double m = 0;
        const int64_t n = sizeX * sizeY * sizeZ;
        for (int64_t i = 0; i < n; i++) m += data[i] / static_cast<double>(n);
        return static_cast<T>(m);
    }
    void set(int64_t x, int64_t y, int64_t z) {
        sizeX = x, sizeY = y, sizeZ = z;
        data = std::make_unique<T[]>(x * y * z);
    }
    // Raw native-endian T without header (volume.h:105-132).
    void load(const std::string& filename, int64_t x, int64_t y, int64_t z) {
        set(x, y, z);
        std::ifstream f(filename, std::ios::binary);
        if (!f) throw std::runtime_error("Volume::load: cannot open " + filename);
        f.read(reinterpret_cast<char*>(data.get()), sizeof(T) * x * y * z);
        if (!f) throw std::runtime_error("Volume::load: short file " + filename);
    }
    void save(const std::string& filename) const {
        std::ofstream f(filename, std::ios::binary);
        if (!f) throw std::runtime_error("Volume::save: cannot open " + filename);
        f.write(reinterpret_cast<const char*>(data.get()), sizeof(T) * sizeX * sizeY * sizeZ);
    }
    void forEach(const std::function<T(T)>& fn) {
        const int64_t n = sizeX * sizeY * sizeZ;
        for (int64_t i = 0; i < n; i++) data[i] = fn(data[i]);
    }
    int64_t x() const { return sizeX; }
    int64_t y() const { return sizeY; }
    int64_t z() const { return sizeZ; }

private:
    int64_t sizeX, sizeY, sizeZ;
    std::unique_ptr<T[]> data;
};

// ---- init_params and its globals (params.cu:14-40) ------------------------
inline std::string PROJ_PATH, VOLUME_PATH, DIRECTION_PATH;
inline int BLOCK_SIZE = 32;
inline int NUM_BASIS_VECTOR = 0;
inline int NUM_PROJ_COND = 1;
inline float SRC_OBJ_DISTANCE = 0, SRC_DETECT_DISTANCE = 0, DETECTOR_SIZE = 0;
inline int NUM_PROJ = 0, NUM_DETECT_U = 0, NUM_DETECT_V = 0, NUM_VOXEL = 0;
inline int LOAD_INDEX[100];

namespace xr_detail {
inline xr_settings& settings() {
    static xr_settings s{};
    return s;
}
inline bool& loaded() {
    static bool b = false;
    return b;
}
inline xr_context* context() {
    static xr_context* ctx = nullptr;
    if (!ctx) {
        const int rc = xr_context_create(&ctx, 0);
        if (rc != XR_OK) throw std::runtime_error("xr_context_create failed (" + std::to_string(rc) + ")");
    }
    return ctx;
}
inline void check(int rc, const char* what) {
    if (rc != XR_OK)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " +
                                 xr_last_error(context()));
}
inline xr_geometry geometry(const Geometry& g, int64_t det_u, int64_t det_v, int64_t n_proj, int64_t n_vox) {
    xr_geometry x{};
    x.sdd = g.sdd;
    x.sod = g.sod;
    x.vox_size = g.voxSize;
    x.det_size = g.detSize;
    x.n_vox = (int32_t)n_vox;
    x.n_det_u = (int32_t)det_u;
    x.n_det_v = (int32_t)det_v;
    x.n_proj = (int32_t)n_proj;
    return x;
}
// Scan tables from init_params, or the identity single-condition scan.
struct ScanTables {
    std::vector<float> R, t, off, basis;
    xr_scan scan{};
};
inline ScanTables scan_tables(int n_cond, int n_basis) {
    ScanTables st;
    if (loaded()) {
        const xr_settings& s = settings();
        st.R.assign(s.cond_R, s.cond_R + 9 * n_cond);
        st.t.assign(s.cond_t, s.cond_t + 3 * n_cond);
        st.off.assign(s.cond_offset, s.cond_offset + 3 * n_cond);
        st.basis.assign(s.basis, s.basis + 3 * n_basis);
    } else {
        for (int c = 0; c < n_cond; c++) {
            const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            st.R.insert(st.R.end(), I, I + 9);
            st.t.insert(st.t.end(), 3, 0.f);
            st.off.insert(st.off.end(), 3, 0.f);
        }
    }
    st.scan.n_cond = n_cond;
    st.scan.n_basis = n_basis;
    st.scan.cond_R = st.R.data();
    st.scan.cond_t = st.t.data();
    st.scan.cond_offset = st.off.data();
    st.scan.basis = st.basis.empty() ? nullptr : st.basis.data();
    st.scan.grating = nullptr;
    return st;
}
// The reference writes the per-epoch loss / (U V P) to ../python/loss.csv,
// relative to the working directory, after every IR / XTT solve
// (reconstruct.cu:153-155, 951-953); an ofstream it cannot open (no such
// directory) writes nothing.  XR_LOSS_CSV names another path; an empty
// XR_LOSS_CSV turns the file off.
inline std::string loss_path() {
    const char* env = std::getenv("XR_LOSS_CSV");
    return env ? std::string(env) : std::string("../python/loss.csv");
}
inline void write_loss(const std::vector<double>& losses, double norm) {
    const std::string path = loss_path();
    if (path.empty()) return;
    std::ofstream f(path);
    for (double l : losses) f << (float)(l / norm) << ",";
}
inline std::string raw_suffix(int64_t a, int64_t b, int64_t c) {
    return "_" + std::to_string(a) + "x" + std::to_string(b) + "x" + std::to_string(c) + ".raw";
}
}  // namespace xr_detail

// ---- raw file names of the reference's driver (main.cu:31-32, 83-107) ------
// Raw float32 files without header, named after init_params' paths and sizes.
namespace xr {
// Sinogram of condition `cond` as loaded: PROJ_PATH + LOAD_INDEX + "_UxVxP.raw".
inline std::string sinogram_load_path(int cond) {
    return PROJ_PATH + std::to_string(LOAD_INDEX[cond]) + xr_detail::raw_suffix(NUM_DETECT_U, NUM_DETECT_V, NUM_PROJ);
}
// Sinogram of condition `cond` as saved: VOLUME_PATH + "_proj" + (cond+1) + "_UxVxP.raw".
inline std::string sinogram_save_path(int cond) {
    return VOLUME_PATH + "_proj" + std::to_string(cond + 1) +
           xr_detail::raw_suffix(NUM_DETECT_U, NUM_DETECT_V, NUM_PROJ);
}
// Volume (coefficient) k: VOLUME_PATH + (k+1) + "_NxNxN.raw".
inline std::string volume_save_path(int k) {
    return VOLUME_PATH + std::to_string(k + 1) + xr_detail::raw_suffix(NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
}
// Main-direction component i (0, 1, 2): DIRECTION_PATH + (i+1) + "_NxNxN.raw".
inline std::string direction_save_path(int i) {
    return DIRECTION_PATH + std::to_string(i + 1) + xr_detail::raw_suffix(NUM_VOXEL, NUM_VOXEL, NUM_VOXEL);
}
// Gradient volume i of IR::gradReconstruct: VOLUME_PATH + "grad" + (i+1) + "_(N+1)^3.raw"
// (reconstruct.cu:316-319).
inline std::string grad_save_path(int i) {
    return VOLUME_PATH + "grad" + std::to_string(i + 1) +
           xr_detail::raw_suffix(NUM_VOXEL + 1, NUM_VOXEL + 1, NUM_VOXEL + 1);
}
// Angle volume i (0: phi, 1: theta): DIRECTION_PATH + "_phi_" / "_theta_" + "NxNxN.raw".
inline std::string angle_save_path(int i) {
    const std::string n = std::to_string(NUM_VOXEL);
    return DIRECTION_PATH + (i == 0 ? "_phi_" : "_theta_") + n + "x" + n + "x" + n + ".raw";
}
}  // namespace xr

inline void init_params(const std::string& tag, const std::string& path) {
    xr_settings& s = xr_detail::settings();
    const int rc = xr_load_settings(path.c_str(), tag.c_str(), &s);
    if (rc != XR_OK) throw std::runtime_error("init_params: cannot load tag '" + tag + "' from " + path);
    PROJ_PATH = s.proj_path;
    VOLUME_PATH = s.vol_path;
    DIRECTION_PATH = s.direc_path;
    BLOCK_SIZE = s.block_size;
    NUM_BASIS_VECTOR = s.n_basis;
    NUM_PROJ_COND = s.n_cond;
    SRC_OBJ_DISTANCE = s.sod;
    SRC_DETECT_DISTANCE = s.sdd;
    NUM_PROJ = s.n_proj;
    NUM_DETECT_U = s.n_det_u;
    NUM_DETECT_V = s.n_det_v;
    DETECTOR_SIZE = s.det_size;
    NUM_VOXEL = s.n_vox;
    for (int i = 0; i < s.n_cond && i < 64; i++) LOAD_INDEX[i] = s.load_index[i];
    xr_detail::loaded() = true;
}

inline void init_params(const std::string& tag) {
    const char* env = std::getenv("XR_SETTINGS");
    init_params(tag, env ? std::string(env) : std::string("settings.json"));
}

// ---- solvers (reconstruct.cuh:26-73) ---------------------------------------
namespace IR {
inline void reconstruct(Volume<float>* sinogram, Volume<float>* voxel, const Geometry& geom, int epoch, int batch,
                        Rotate dir, Method method, float lambda = 1e-2) {
    const int C = NUM_PROJ_COND;
    auto st = xr_detail::scan_tables(C, 0);
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), voxel[0].x());
    std::vector<float*> s(C), v(C);
    for (int c = 0; c < C; c++) {
        s[c] = sinogram[c].get();
        v[c] = voxel[c].get();
    }
    std::vector<double> losses((size_t)C * epoch);
    xr_stats stats{losses.data()};
    xr_detail::check(xr_ir_reconstruct(xr_detail::context(), &g, &st.scan, s.data(), v.data(), epoch, batch,
                                       dir == Rotate::CW ? XR_CW : XR_CCW,
                                       method == Method::ART ? XR_METHOD_ART : XR_METHOD_MLEM, lambda, &stats),
                     "IR::reconstruct");
    // loss.csv holds the last condition's epochs (the reference rewrites it per call)
    std::vector<double> last(losses.end() - epoch, losses.end());
    xr_detail::write_loss(last, (double)g.n_det_u * g.n_det_v * g.n_proj);
}

// IR::gradReconstruct (reconstruct.cu:158-328): gradient-domain phase contrast
// from sinogram[0].  Like the reference it advances geom.voxel by one, writes
// the Poisson-integrated volume to voxel[0] and saves the gradient volumes to
// grad_save_path(i) for i < 3*min(NUM_BASIS_VECTOR, NUM_PROJ_COND) (the
// reference's loop runs to 3*NUM_BASIS_VECTOR over 3*NUM_PROJ_COND volumes).
// The launch extent follows BLOCK_SIZE as the reference's does.
inline void gradReconstruct(Volume<float>* sinogram, Volume<float>* voxel, Geometry& geom, int epoch, int batch,
                            Rotate dir, Method method, float lambda) {
    geom.voxel = geom.voxel + 1;
    const int C = NUM_PROJ_COND;
    auto st = xr_detail::scan_tables(C, 0);
    const int64_t N = voxel[0].x();
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), N);
    std::vector<Volume<float>> grad(3 * C);
    std::vector<float*> gp(3 * C);
    for (int i = 0; i < 3 * C; i++) {
        grad[i] = Volume<float>(N + 1, N + 1, N + 1);
        gp[i] = grad[i].get();
    }
    xr_detail::check(xr_grad_reconstruct(xr_detail::context(), &g, &st.scan, sinogram[0].get(), voxel[0].get(),
                                         gp.data(), epoch, batch, dir == Rotate::CW ? XR_CW : XR_CCW,
                                         method == Method::ART ? XR_METHOD_ART : XR_METHOD_MLEM, lambda, BLOCK_SIZE,
                                         nullptr),
                     "IR::gradReconstruct");
    const int n_save = 3 * std::min(NUM_BASIS_VECTOR, C);
    for (int i = 0; i < n_save; i++) grad[i].save(xr::grad_save_path(i));
}
}  // namespace IR

namespace XTT {
inline void reconstruct(Volume<float>* sinogram, Volume<float>* voxel, Volume<float>* md, const Geometry& geom,
                        int epoch, int batch, Rotate dir, Method method, float lambda = 1e-2) {
    const int C = NUM_PROJ_COND, K = NUM_BASIS_VECTOR;
    if (K <= 0) throw std::runtime_error("XTT::reconstruct: init_params must set basis vectors");
    auto st = xr_detail::scan_tables(C, K);
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), voxel[0].x());
    std::vector<float*> s(C), v(K), m(3);
    for (int c = 0; c < C; c++) s[c] = sinogram[c].get();
    for (int k = 0; k < K; k++) v[k] = voxel[k].get();
    for (int i = 0; i < 3; i++) {
        if (md[i].x() != voxel[0].x()) md[i] = Volume<float>(voxel[0].x(), voxel[0].y(), voxel[0].z());
        m[i] = md[i].get();
    }
    std::vector<double> losses(epoch);
    xr_stats stats{losses.data()};
    xr_detail::check(xr_xtt_reconstruct(xr_detail::context(), &g, &st.scan, s.data(), v.data(), m.data(), nullptr,
                                        epoch, batch, dir == Rotate::CW ? XR_CW : XR_CCW,
                                        method == Method::ART ? XR_METHOD_ART : XR_METHOD_MLEM, lambda, &stats),
                     "XTT::reconstruct");
    xr_detail::write_loss(losses, (double)g.n_det_u * g.n_det_v * g.n_proj);
}

// XTT::orthTwiceReconstruct (reconstruct.cu:956-1239): the orientation-aware
// three-direction model.  voxel[3] is the initial guess in and the kept volume
// out, md[3] the main directions; per outer iteration the directions and
// volumes are written to DIRECTION_PATH / VOLUME_PATH + "_sequence<outer>d<i>
// _NxNxN.raw" as the reference does.  The subset order is shuffled by a
// std::mt19937 seeded from std::random_device, as in the reference (pass a
// seed to xr_orth_reconstruct for a reproducible run).
inline void orthTwiceReconstruct(Volume<float>* sinogram, Volume<float> voxel[3], Volume<float> md[3],
                                 const Geometry& geom, int iter1, int iter2, int batch, Rotate dir, Method method,
                                 float lambda) {
    const int C = NUM_PROJ_COND;
    if (NUM_BASIS_VECTOR != 3) throw std::runtime_error("XTT::orthTwiceReconstruct: needs three basis vectors");
    auto st = xr_detail::scan_tables(C, 3);
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), voxel[0].x());
    std::vector<float*> s(C), v(3), m(3);
    for (int c = 0; c < C; c++) s[c] = sinogram[c].get();
    for (int k = 0; k < 3; k++) {
        v[k] = voxel[k].get();
        if (md[k].x() != voxel[0].x()) md[k] = Volume<float>(voxel[0].x(), voxel[0].y(), voxel[0].z());
        m[k] = md[k].get();
    }
    std::random_device seed_gen;
    xr_detail::check(xr_orth_reconstruct(xr_detail::context(), &g, &st.scan, s.data(), v.data(), m.data(), iter1,
                                         iter2, batch, dir == Rotate::CW ? XR_CW : XR_CCW,
                                         method == Method::ART ? XR_METHOD_ART : XR_METHOD_MLEM, lambda, seed_gen(),
                                         VOLUME_PATH.c_str(), DIRECTION_PATH.c_str()),
                     "XTT::orthTwiceReconstruct");
}
}  // namespace XTT

namespace FDK {
inline void reconstruct(Volume<float>* sinogram, Volume<float>* voxel, const Geometry& geom, Rotate dir) {
    const int C = NUM_PROJ_COND;
    auto st = xr_detail::scan_tables(C, 0);
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), voxel[0].x());
    std::vector<float*> s(C), v(C);
    for (int c = 0; c < C; c++) {
        s[c] = sinogram[c].get();
        v[c] = voxel[c].get();
    }
    xr_detail::check(xr_fdk_reconstruct(xr_detail::context(), &g, &st.scan, s.data(), v.data(),
                                        dir == Rotate::CW ? XR_CW : XR_CCW, XR_FILTER_REFERENCE),
                     "FDK::reconstruct");
}
// FDK::hilbertReconstruct (reconstruct.cu:1726-1793): every condition's raw
// rows through the Hilbert filter of fdk.cu:298-375 (XR_FILTER_HILBERT), then
// the FDK weighted backprojection.  The filtered rows are returned through the
// sinograms.  The reference filters only condition 0 (the others are
// backprojected unfiltered) and copies into a device buffer it never
// allocates (reconstruct.cu:1749); here every condition is filtered.
inline void hilbertReconstruct(Volume<float>* sinogram, Volume<float>* voxel, const Geometry& geom, Rotate dir) {
    const int C = NUM_PROJ_COND;
    auto st = xr_detail::scan_tables(C, 0);
    xr_geometry g = xr_detail::geometry(geom, sinogram[0].x(), sinogram[0].y(), sinogram[0].z(), voxel[0].x());
    std::vector<float*> s(C), v(C);
    for (int c = 0; c < C; c++) {
        s[c] = sinogram[c].get();
        v[c] = voxel[c].get();
    }
    xr_detail::check(xr_fdk_reconstruct(xr_detail::context(), &g, &st.scan, s.data(), v.data(),
                                        dir == Rotate::CW ? XR_CW : XR_CCW, XR_FILTER_HILBERT),
                     "FDK::hilbertReconstruct");
}
}  // namespace FDK

// ---- direction utilities (direction/include/pca.cuh:9-12) ------------------

// calcEigenVector for one voxel (pca.cu:12-65), on the GPU.  For whole volumes
// use calcEigenVectorVolume (one launch instead of one per voxel).
inline void calcEigenVector(const Volume<float>* ct, Volume<float>* md, Volume<float>* evalue, int x, int y, int z) {
    const int K = NUM_BASIS_VECTOR;
    if (K <= 0) throw std::runtime_error("calcEigenVector: init_params must set basis vectors");
    std::vector<float> mu(K), out(6);
    for (int k = 0; k < K; k++) mu[k] = ct[k](x, y, z);
    xr_detail::check(xr_pca_host(xr_detail::context(), mu.data(), K, xr_detail::settings().basis, 1, out.data(),
                                 out.data() + 3),
                     "calcEigenVector");
    for (int i = 0; i < 3; i++) {
        md[i](x, y, z) = out[i];
        evalue[i](x, y, z) = out[3 + i];
    }
}

// calcEigenVector over every voxel (what XTT::reconstruct runs after its
// iterations, reconstruct.cu:935-942).  md/evalue are resized to ct's shape.
inline void calcEigenVectorVolume(const Volume<float>* ct, Volume<float>* md, Volume<float>* evalue) {
    const int K = NUM_BASIS_VECTOR;
    if (K <= 0) throw std::runtime_error("calcEigenVectorVolume: init_params must set basis vectors");
    const int64_t n = ct[0].x() * ct[0].y() * ct[0].z();
    std::vector<float> mu((size_t)K * n), out((size_t)6 * n);
    for (int k = 0; k < K; k++) std::memcpy(mu.data() + (size_t)k * n, ct[k].get(), sizeof(float) * n);
    xr_detail::check(xr_pca_host(xr_detail::context(), mu.data(), K, xr_detail::settings().basis, n, out.data(),
                                 out.data() + 3 * n),
                     "calcEigenVectorVolume");
    for (int i = 0; i < 3; i++) {
        md[i] = Volume<float>(ct[0].x(), ct[0].y(), ct[0].z());
        evalue[i] = Volume<float>(ct[0].x(), ct[0].y(), ct[0].z());
        std::memcpy(md[i].get(), out.data() + (size_t)i * n, sizeof(float) * n);
        std::memcpy(evalue[i].get(), out.data() + (size_t)(3 + i) * n, sizeof(float) * n);
    }
}

// calcAngleFromMD (pca.cu:67-83): azimuth/polar angles of the main direction
// (computed on the GPU, xr_md_angles).
inline void calcAngleFromMD(const Volume<float> md[3], Volume<float> angle[2], int size_x, int size_y, int size_z) {
    const int64_t n = (int64_t)size_x * size_y * size_z;
    if (md[0].x() * md[0].y() * md[0].z() != n)
        throw std::runtime_error("calcAngleFromMD: sizes do not match the direction volumes");
    std::vector<float> in((size_t)3 * n), out((size_t)2 * n);
    for (int i = 0; i < 3; i++) std::memcpy(in.data() + (size_t)i * n, md[i].get(), sizeof(float) * n);
    xr_detail::check(xr_md_angles_host(xr_detail::context(), in.data(), out.data(), n), "calcAngleFromMD");
    for (int i = 0; i < 2; i++) {
        if (angle[i].x() * angle[i].y() * angle[i].z() != n) angle[i] = Volume<float>(size_x, size_y, size_z);
        std::memcpy(angle[i].get(), out.data() + (size_t)i * n, sizeof(float) * n);
    }
}

// directionNormalization (pca.cu:85-106): unit main direction where |md| > 0.65,
// zero elsewhere (computed on the GPU, xr_md_normalise).
inline void directionNormalization(Volume<float> md[3]) {
    const int64_t n = md[0].x() * md[0].y() * md[0].z();
    std::vector<float> buf((size_t)3 * n);
    for (int i = 0; i < 3; i++) std::memcpy(buf.data() + (size_t)i * n, md[i].get(), sizeof(float) * n);
    xr_detail::check(xr_md_normalise_host(xr_detail::context(), buf.data(), n, 0.65f), "directionNormalization");
    for (int i = 0; i < 3; i++) std::memcpy(md[i].get(), buf.data() + (size_t)i * n, sizeof(float) * n);
}
