// Text formats of the reference: scene directories (objectScene::saveSelf /
// initFromFile, SceneUtils.cu:1246-1299, 703-790; file helpers Utils.cu:
// 554-673), camera files (Camera::readFromMyBlender, CameraUtils.cu:
// 271-386) and image text (Image::initSelfFromFile + readMask,
// ImageUtils.cu:84-128, Utils.cu:508-572; written by
// utils/preprocess_datasets.py:9-22).  Values are written with the
// reference's format strings ("%.6lf " floats, "%d " integers, one row per
// line) and read back as whitespace-separated numbers.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace pxl {
namespace {

struct TextFile {
    FILE *f;
    TextFile(const std::string &p, const char *mode) : f(std::fopen(p.c_str(), mode)) {}
    ~TextFile() {
        if (f) std::fclose(f);
    }
};

std::string join(const char *dir, const std::string &name) {
    std::string s(dir);
    if (!s.empty() && s.back() != '/') s += '/';
    return s + name;
}

std::string fmt(const char *f, int a) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), f, a);
    return buf;
}

// rows x cols values of get(i), "%.6lf " each, a newline per row
// (saveFloaters2File, Utils.cu:648-673).
template <typename Get>
bool write_floats(const std::string &p, int64_t rows, int64_t cols, Get get) {
    TextFile t(p, "w");
    if (!t.f) return false;
    for (int64_t r = 0; r < rows; r++) {
        for (int64_t c = 0; c < cols; c++) std::fprintf(t.f, "%.6lf ", (double)get(r * cols + c));
        std::fprintf(t.f, "\n");
    }
    return true;
}

// "%d " integers (saveUchar2File / saveInt2File, Utils.cu:575-622).
template <typename Get>
bool write_ints(const std::string &p, int64_t rows, int64_t cols, Get get) {
    TextFile t(p, "w");
    if (!t.f) return false;
    for (int64_t r = 0; r < rows; r++) {
        for (int64_t c = 0; c < cols; c++) std::fprintf(t.f, "%d ", (int)get(r * cols + c));
        std::fprintf(t.f, "\n");
    }
    return true;
}

// n whitespace-separated floats ("%f", readFloatersFromFile, Utils.cu:624-646).
bool read_floats(const std::string &p, int64_t n, std::vector<float> &out) {
    TextFile t(p, "r");
    if (!t.f) return false;
    out.resize((size_t)n);
    for (int64_t i = 0; i < n; i++)
        if (std::fscanf(t.f, "%f", &out[(size_t)i]) != 1) return false;
    return true;
}

// Up to n integers ("%d"); returns how many were read, -1 if no file.
int64_t read_ints(const std::string &p, int64_t n, std::vector<int> &out) {
    TextFile t(p, "r");
    if (!t.f) return -1;
    out.resize((size_t)n);
    int64_t i = 0;
    while (i < n && std::fscanf(t.f, "%d", &out[(size_t)i]) == 1) i++;
    return i;
}

float half_to_float(uint16_t h) {
    const uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 0x1F, m = h & 0x3FF;
    uint32_t bits;
    if (e == 0) {
        if (m == 0) {
            bits = s;
        } else {  // subnormal: normalise
            int sh = 0;
            uint32_t mm = m;
            while (!(mm & 0x400)) {
                mm <<= 1;
                sh++;
            }
            bits = s | ((uint32_t)(113 - sh) << 23) | ((mm & 0x3FF) << 13);
        }
    } else if (e == 31) {
        bits = s | 0x7F800000u | (m << 13);
    } else {
        bits = s | ((e + 112) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}

// In-cage flags of the occupied cells (pruneScene, SceneUtils.cu:836-878),
// including the face rule that marks H*W + x + 1 for x = W - 2.
void cage_from_occupancy(const std::vector<uint8_t> &occ, int W, int H, int D,
                         std::vector<uint8_t> &cage) {
    cage.assign((size_t)W * H * D, 0);
    const int64_t WH = (int64_t)W * H;
    for (int z = 0; z < D - 1; z++)
        for (int y = 0; y < H - 1; y++)
            for (int x = 0; x < W - 1; x++) {
                if (!occ[((size_t)z * (H - 1) + y) * (W - 1) + x]) continue;
                cage[(size_t)(z * WH + (int64_t)y * W + x)] = 1;
                if (z == D - 2) cage[(size_t)((D - 1) * WH + (int64_t)y * W + x)] = 1;
                if (y == H - 2) cage[(size_t)(z * WH + (int64_t)(H - 1) * W + x)] = 1;
                if (x == W - 2) cage[(size_t)(z * WH + WH + x + 1)] = 1;
            }
}

int32_t d2h(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return PXL_OK;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("device to host copy of %zu bytes failed", bytes);
        return PXL_EHIP;
    }
    return PXL_OK;
}

int32_t h2d(void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return PXL_OK;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        set_error("host to device copy of %zu bytes failed", bytes);
        return PXL_EHIP;
    }
    return PXL_OK;
}

const char *kChannel[3] = {"red", "green", "blue"};

}  // namespace
}  // namespace pxl

using namespace pxl;

extern "C" int32_t pxl_read_cameras(const char *path, const double object_centre[3],
                                    int32_t target_w, int32_t target_h, int32_t camera_nums,
                                    double camera_distance, int32_t need_rescale,
                                    int32_t z_reversed, pxl_camera *out, int32_t capacity,
                                    int32_t *count) {
    clear_error();
    PXL_REQUIRE(path && object_centre && count, "pxl_read_cameras: null argument");
    TextFile t(path, "r");
    PXL_REQUIRE(t.f, "pxl_read_cameras: cannot open %s", path);
    int n = 0;
    PXL_REQUIRE(std::fscanf(t.f, " total %d", &n) == 1 && n >= 0,
                "pxl_read_cameras: %s: no 'total' line", path);
    if (camera_nums != 0) n = camera_nums;
    double K[9];
    for (int i = 0; i < 9; i++)
        PXL_REQUIRE(std::fscanf(t.f, "%lf", &K[i]) == 1, "pxl_read_cameras: %s: short intrinsics",
                    path);
    // Rescale the intrinsics when the principal point is not the centre of
    // the target image (CameraUtils.cu:304-317): K <- diag(sx, sy, 1) K.
    double S[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    if (std::fabs(K[2] * 2 - target_w) > 1.e-8) S[0] = target_w / K[2] / 2;
    if (std::fabs(K[5] * 2 - target_h) > 1.e-8) S[4] = target_h / K[5] / 2;
    double Ks[9];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += S[r * 3 + k] * K[k * 3 + c];
            Ks[r * 3 + c] = s;
        }
    *count = n;
    if (!out) return PXL_OK;
    PXL_REQUIRE(capacity >= n, "pxl_read_cameras: %d cameras, room for %d", n, capacity);
    for (int i = 0; i < n; i++) {
        int idx = -1;
        PXL_REQUIRE(std::fscanf(t.f, " %d:", &idx) == 1 && idx == i,
                    "pxl_read_cameras: %s: view %d header missing or out of order", path, i);
        double tr[3], w2c[16];
        for (int k = 0; k < 3; k++)
            PXL_REQUIRE(std::fscanf(t.f, "%lf", &tr[k]) == 1,
                        "pxl_read_cameras: %s: view %d translation", path, i);
        // The stored world-to-camera matrix is read and not used: the
        // reference rebuilds the rotation by look-at (CameraUtils.cu:371-377).
        for (int k = 0; k < 16; k++)
            PXL_REQUIRE(std::fscanf(t.f, "%lf", &w2c[k]) == 1,
                        "pxl_read_cameras: %s: view %d matrix", path, i);
        const int32_t st = pxl_camera_from_pose(tr, object_centre, camera_distance, need_rescale,
                                                z_reversed, Ks[0], Ks[4], Ks[2], Ks[5], &out[i],
                                                nullptr);
        if (st != PXL_OK) {
            set_error("pxl_read_cameras: view %d: singular camera", i);
            return st;
        }
    }
    return PXL_OK;
}

extern "C" int32_t pxl_read_image_text(const char *prefix, int32_t width, int32_t height,
                                       int32_t gray_only, uint8_t *r, uint8_t *g, uint8_t *b,
                                       uint8_t *mask, int32_t bounds[4]) {
    clear_error();
    PXL_REQUIRE(prefix && r && mask && bounds && (gray_only || (g && b)),
                "pxl_read_image_text: null argument");
    PXL_REQUIRE(width > 0 && height > 0, "pxl_read_image_text: bad size %dx%d", width, height);
    const int64_t n = (int64_t)width * height;
    std::vector<int> v;
    const std::string mp = std::string(prefix) + "_mask.txt";
    PXL_REQUIRE(read_ints(mp, n, v) == n, "pxl_read_image_text: cannot read %lld values of %s",
                (long long)n, mp.c_str());
    bounds[0] = height - 1;
    bounds[1] = 0;
    bounds[2] = width - 1;
    bounds[3] = 0;
    for (int64_t i = 0; i < n; i++) {
        mask[i] = (uint8_t)v[(size_t)i];
        if (v[(size_t)i]) {
            const int row = (int)(i / width), col = (int)(i % width);
            if (row < bounds[0]) bounds[0] = row;
            if (row > bounds[1]) bounds[1] = row;
            if (col < bounds[2]) bounds[2] = col;
            if (col > bounds[3]) bounds[3] = col;
        }
    }
    const char *sfx[3] = {"_red.txt", "_green.txt", "_blue.txt"};
    uint8_t *dst[3] = {r, g, b};
    for (int c = 0; c < (gray_only ? 1 : 3); c++) {
        const std::string p = std::string(prefix) + (gray_only ? ".txt" : sfx[c]);
        PXL_REQUIRE(read_ints(p, n, v) == n, "pxl_read_image_text: cannot read %lld values of %s",
                    (long long)n, p.c_str());
        for (int64_t i = 0; i < n; i++) dst[c][i] = (uint8_t)v[(size_t)i];
    }
    return PXL_OK;
}

extern "C" int32_t pxl_grid_save_text(const pxl_grid *g, const char *dir, int32_t band,
                                      int32_t density_only, int32_t hierarchy_mode,
                                      void *stream) {
    clear_error();
    PXL_REQUIRE(g && dir && g->params, "pxl_grid_save_text: null argument");
    PXL_REQUIRE(g->W >= 2 && g->H >= 2 && g->D >= 2, "pxl_grid_save_text: bad dims");
    const int nsh = (band + 1) * (band + 1);
    PXL_REQUIRE(band >= 0 && nsh <= g->sh_dim, "pxl_grid_save_text: band %d needs %d of %d "
                "SH coefficients", band, nsh, g->sh_dim);
    PXL_REQUIRE(!g->sparse || (g->in_cage && g->index_offsets),
                "pxl_grid_save_text: a caged grid needs in_cage and index_offsets");
    const hipStream_t st = as_stream(stream);
    const int W = g->W, H = g->H, D = g->D, S = g->sh_dim, stride = g->stride;
    const int64_t nv = (int64_t)W * H * D, ncell = (int64_t)(W - 1) * (H - 1) * (D - 1);
    const int64_t n = g->n_vox;
    const size_t esz = g->storage == PXL_STORE_F32 ? 4 : 2;
    std::vector<uint8_t> raw((size_t)n * stride * esz);
    int32_t rc = d2h(raw.data(), g->params, raw.size(), st);
    if (rc != PXL_OK) return rc;
    std::vector<float> rec((size_t)n * stride);
    for (size_t i = 0; i < rec.size(); i++) {
        if (esz == 4) {
            std::memcpy(&rec[i], &raw[i * 4], 4);
        } else {
            uint16_t h;
            std::memcpy(&h, &raw[i * 2], 2);
            if (g->storage == PXL_STORE_BF16) {
                const uint32_t bits = (uint32_t)h << 16;
                std::memcpy(&rec[i], &bits, 4);
            } else {
                rec[i] = half_to_float(h);
            }
        }
    }
    // Occupancy bytes (0/1) per cell; a grid without a bitfield saves zeros.
    std::vector<uint32_t> bits((size_t)((ncell + 31) / 32), 0u);
    if (g->occupancy && (rc = d2h(bits.data(), g->occupancy, bits.size() * 4, st)) != PXL_OK)
        return rc;
    bool ok = write_ints(join(dir, "occupancy.txt"), D - 1, (int64_t)(W - 1) * (H - 1),
                         [&](int64_t c) { return (bits[(size_t)(c >> 5)] >> (c & 31)) & 1u; });
    auto val = [&](int q) { return [&, q](int64_t r) { return rec[(size_t)r * stride + q]; }; };
    if (!g->sparse) {
        ok = ok && write_floats(join(dir, fmt("opacity_band%d.txt", band)), D, (int64_t)W * H,
                                val(0));
    } else {
        std::vector<uint8_t> cage((size_t)nv);
        std::vector<uint32_t> io((size_t)nv);
        if ((rc = d2h(cage.data(), g->in_cage, cage.size(), st)) != PXL_OK ||
            (rc = d2h(io.data(), g->index_offsets, io.size() * 4, st)) != PXL_OK)
            return rc;
        ok = ok && write_floats(join(dir, fmt("sparse_opacity_band%d.txt", band)), n, 1, val(0));
        ok = ok && write_ints(join(dir, fmt("sparse_inCageInfo_band%d.txt", band)), D,
                              (int64_t)W * H, [&](int64_t i) { return cage[(size_t)i]; });
        ok = ok && write_ints(join(dir, "sparse_indexOffset.txt"), D, (int64_t)W * H,
                              [&](int64_t i) { return io[(size_t)i]; });
        // The count goes to the file name the reference builds for it (its
        // saveSelf opens the in-cage file again instead, SceneUtils.cu:1279-1283).
        TextFile t(join(dir, "InCageNumsInfo.txt"), "w");
        ok = ok && t.f && std::fprintf(t.f, "%d", (int)n) > 0;
    }
    if (!density_only) {
        // hierarchy_mode saves only the coefficients of the last band
        // (saveChannelSHs, SceneUtils.cu:1207-1244).
        const int first = hierarchy_mode ? band * band : 0;
        for (int ch = 0; ch < 3; ch++)
            for (int i = first; i < nsh; i++) {
                const int q = 1 + ch * S + i;
                if (!g->sparse)
                    ok = ok && write_floats(join(dir, std::string(kChannel[ch]) + fmt("_%d.txt", i)),
                                            D, (int64_t)W * H, val(q));
                else
                    ok = ok && write_floats(
                                   join(dir, std::string(kChannel[ch]) + fmt("_%d_sparse.txt", i)),
                                   n, 1, val(q));
            }
    }
    if (!ok) {
        set_error("pxl_grid_save_text: cannot write into %s", dir);
        return PXL_EINVAL;
    }
    return PXL_OK;
}

extern "C" int32_t pxl_grid_text_records(const char *dir, int32_t band, int32_t sparse,
                                         int32_t W, int32_t H, int32_t D, int64_t *n_records) {
    clear_error();
    PXL_REQUIRE(dir && n_records, "pxl_grid_text_records: null argument");
    if (!sparse) {
        *n_records = (int64_t)W * H * D;
        return PXL_OK;
    }
    std::vector<int> v;
    if (read_ints(join(dir, "InCageNumsInfo.txt"), 1, v) == 1 ||
        read_ints(join(dir, fmt("sparse_inCageInfo_band%d.txt", band)), 1, v) == 1) {
        *n_records = v[0];
        return PXL_OK;
    }
    set_error("pxl_grid_text_records: no in-cage count in %s", dir);
    return PXL_EINVAL;
}

extern "C" int32_t pxl_grid_load_text(const char *dir, int32_t band, int32_t sparse,
                                      int32_t have_occupancy, int32_t W, int32_t H, int32_t D,
                                      int32_t sh_dim, int32_t stride, int64_t n_records,
                                      float *params, uint32_t *occupancy, uint8_t *in_cage,
                                      uint32_t *index_offsets, uint32_t *compact_to_dense,
                                      void *stream) {
    clear_error();
    const int nsh = (band + 1) * (band + 1);
    PXL_REQUIRE(dir && params, "pxl_grid_load_text: null argument");
    PXL_REQUIRE(W >= 2 && H >= 2 && D >= 2 && band >= 0 && nsh <= sh_dim &&
                    stride >= 1 + 3 * sh_dim,
                "pxl_grid_load_text: bad dims / band %d / sh_dim %d / stride %d", band, sh_dim,
                stride);
    PXL_REQUIRE(!have_occupancy || occupancy, "pxl_grid_load_text: null occupancy output");
    PXL_REQUIRE(!sparse || (in_cage && index_offsets && compact_to_dense && have_occupancy),
                "pxl_grid_load_text: a caged grid needs occupancy, in_cage, index_offsets and "
                "compact_to_dense outputs");
    const hipStream_t st = as_stream(stream);
    const int64_t nv = (int64_t)W * H * D, ncell = (int64_t)(W - 1) * (H - 1) * (D - 1);
    const int64_t n = sparse ? n_records : nv;
    PXL_REQUIRE(n >= 1 && n <= nv, "pxl_grid_load_text: %lld records", (long long)n);
    // Everything is read and checked on the host before the first copy to
    // the device, so a rejected directory leaves the device buffers as they
    // were.
    std::vector<int> iv;
    std::vector<uint8_t> occ;
    std::vector<uint32_t> bits;
    if (have_occupancy) {
        PXL_REQUIRE(read_ints(join(dir, "occupancy.txt"), ncell, iv) == ncell,
                    "pxl_grid_load_text: occupancy.txt of %s is short", dir);
        occ.resize((size_t)ncell);
        bits.assign((size_t)((ncell + 31) / 32), 0u);
        for (int64_t c = 0; c < ncell; c++) {
            occ[(size_t)c] = (uint8_t)iv[(size_t)c];
            if (occ[(size_t)c]) bits[(size_t)(c >> 5)] |= 1u << (c & 31);
        }
    }
    std::vector<float> rec((size_t)n * stride, 0.0f), fv;
    const std::string dens = sparse ? fmt("sparse_opacity_band%d.txt", band)
                                    : fmt("opacity_band%d.txt", band);
    PXL_REQUIRE(read_floats(join(dir, dens), n, fv), "pxl_grid_load_text: cannot read %s",
                dens.c_str());
    for (int64_t r = 0; r < n; r++) rec[(size_t)r * stride] = fv[(size_t)r];
    for (int ch = 0; ch < 3; ch++)
        for (int i = 0; i < nsh; i++) {
            const std::string f = std::string(kChannel[ch]) +
                                  (sparse ? fmt("_%d_sparse.txt", i) : fmt("_%d.txt", i));
            PXL_REQUIRE(read_floats(join(dir, f), n, fv), "pxl_grid_load_text: cannot read %s",
                        f.c_str());
            for (int64_t r = 0; r < n; r++)
                rec[(size_t)r * stride + 1 + ch * sh_dim + i] = fv[(size_t)r];
        }
    std::vector<uint8_t> cage;
    std::vector<uint32_t> io, c2d;
    if (sparse) {
        // In-cage flags from their file when it holds them; a directory
        // the reference saved holds only the count there, and the flags
        // follow from the occupancy by the pruning rule.
        const int64_t got = read_ints(join(dir, fmt("sparse_inCageInfo_band%d.txt", band)), nv, iv);
        if (got == nv) {
            cage.resize((size_t)nv);
            for (int64_t i = 0; i < nv; i++) cage[(size_t)i] = (uint8_t)iv[(size_t)i];
        } else {
            cage_from_occupancy(occ, W, H, D, cage);
        }
        PXL_REQUIRE(read_ints(join(dir, "sparse_indexOffset.txt"), nv, iv) == nv,
                    "pxl_grid_load_text: sparse_indexOffset.txt of %s is short", dir);
        io.resize((size_t)nv);
        c2d.reserve((size_t)n);
        for (int64_t i = 0; i < nv; i++) {
            io[(size_t)i] = (uint32_t)iv[(size_t)i];
            if (cage[(size_t)i]) c2d.push_back((uint32_t)i);
        }
        PXL_REQUIRE((int64_t)c2d.size() == n,
                    "pxl_grid_load_text: %zu in-cage vertices, %lld records", c2d.size(),
                    (long long)n);
        for (int64_t k = 1; k < n; k++)
            PXL_REQUIRE(io[c2d[(size_t)k]] == (uint32_t)k,
                        "pxl_grid_load_text: index of vertex %u is %u, expected %lld", c2d[(size_t)k],
                        io[c2d[(size_t)k]], (long long)k);
        // Every other vertex (outside the cage, and the first one in it) has
        // index 0 (getCagedNums, SceneUtils.cu:1106-1113); a file that says
        // otherwise would send kernels to records that do not exist.
        for (int64_t i = 0; i < nv; i++)
            PXL_REQUIRE(cage[(size_t)i] || io[(size_t)i] == 0u,
                        "pxl_grid_load_text: vertex %lld is outside the cage but has index %u",
                        (long long)i, io[(size_t)i]);
        PXL_REQUIRE(n == 0 || io[c2d[0]] == 0u,
                    "pxl_grid_load_text: the first in-cage vertex has index %u, expected 0",
                    io[c2d[0]]);
    }
    int32_t rc;
    if (have_occupancy && (rc = h2d(occupancy, bits.data(), bits.size() * 4, st)) != PXL_OK)
        return rc;
    if (sparse &&
        ((rc = h2d(in_cage, cage.data(), cage.size(), st)) != PXL_OK ||
         (rc = h2d(index_offsets, io.data(), io.size() * 4, st)) != PXL_OK ||
         (rc = h2d(compact_to_dense, c2d.data(), c2d.size() * 4, st)) != PXL_OK))
        return rc;
    return h2d(params, rec.data(), rec.size() * 4, st);
}
