// Sanitizer driver (test infrastructure): the host layer's text parsers,
// camera and shuffle code and the oracle's C restatement, built with
// AddressSanitizer and UndefinedBehaviorSanitizer and run on the CPU over
// valid and corrupt inputs.  Exit status 0 and "san: ok" when every
// expectation holds; a sanitizer report aborts the run.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pxl.h"
#include "oracle.h"

static int g_fail = 0;
#define EXPECT(c)                                                        \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "san: expectation failed at line %d: %s\n", __LINE__, #c); \
            g_fail++;                                                    \
        }                                                                \
    } while (0)

static std::string g_dir;
static std::string path(const std::string &name) { return g_dir + "/" + name; }
static void put(const std::string &name, const std::string &text) {
    FILE *f = std::fopen(path(name).c_str(), "w");
    if (!f) {
        std::perror("san: write");
        std::exit(2);
    }
    std::fputs(text.c_str(), f);
    std::fclose(f);
}
static std::string ints(const std::vector<int> &v) {
    std::string s;
    for (int x : v) s += std::to_string(x) + " ";
    return s + "\n";
}
static std::string floats(const std::vector<float> &v) {
    std::string s;
    char b[64];
    for (float x : v) {
        std::snprintf(b, sizeof b, "%.6lf ", (double)x);
        s += b;
    }
    return s + "\n";
}

static void cameras() {
    const std::string good =
        "total 3\n400 0 200\n0 400 150\n0 0 1\n"
        "0: 1 2 3\n1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1\n"
        "1: -2 1 0.5\n1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1\n"
        "2: 0.1 -3 2\n1 0 0 0 0 1 0 0 0 0 1 0 0 0 0 1\n";
    put("cams.txt", good);
    const double centre[3] = {0.0, 0.0, 0.0};
    pxl_camera cams[4];
    int32_t n = 0;
    EXPECT(pxl_read_cameras(path("cams.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams, 4,
                            &n) == PXL_OK && n == 3);
    EXPECT(pxl_read_cameras(path("cams.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams, 2,
                            &n) == PXL_EINVAL);  // room for 2 of 3
    // Truncated in the middle of a matrix, of the intrinsics, or of the
    // header; a view out of order; a count far beyond the file.
    put("cams_cut.txt", good.substr(0, good.size() / 2));
    EXPECT(pxl_read_cameras(path("cams_cut.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams,
                            4, &n) == PXL_EINVAL);
    put("cams_k.txt", "total 3\n400 0\n");
    EXPECT(pxl_read_cameras(path("cams_k.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams, 4,
                            &n) == PXL_EINVAL);
    put("cams_none.txt", "");
    EXPECT(pxl_read_cameras(path("cams_none.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams,
                            4, &n) == PXL_EINVAL);
    std::string order = good;
    order.replace(order.find("1: "), 3, "7: ");
    put("cams_order.txt", order);
    EXPECT(pxl_read_cameras(path("cams_order.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams,
                            4, &n) == PXL_EINVAL);
    put("cams_big.txt", "total 2000000000\n400 0 200\n0 400 150\n0 0 1\n");
    EXPECT(pxl_read_cameras(path("cams_big.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams, 4,
                            &n) == PXL_EINVAL);
    EXPECT(pxl_read_cameras(path("missing.txt").c_str(), centre, 400, 300, 0, 4.0, 1, 0, cams, 4,
                            &n) == PXL_EINVAL);
    // Degenerate pose: the camera at the object centre.
    const double at[3] = {0.0, 0.0, 0.0};
    pxl_camera c;
    (void)pxl_camera_from_pose(at, centre, 4.0, 0, 0, 400, 400, 200, 150, &c, nullptr);
}

static void images() {
    const int W = 7, H = 5;
    std::vector<int> mask(W * H, 0), red(W * H), green(W * H), blue(W * H);
    mask[2 * W + 3] = 255;
    mask[4 * W + 6] = 128;
    for (int i = 0; i < W * H; i++) {
        red[i] = i % 256;
        green[i] = (3 * i) % 256;
        blue[i] = (7 * i) % 256;
    }
    put("im_mask.txt", ints(mask));
    put("im_red.txt", ints(red));
    put("im_green.txt", ints(green));
    put("im_blue.txt", ints(blue));
    put("im.txt", ints(red));
    std::vector<uint8_t> r(W * H), g(W * H), b(W * H), m(W * H);
    int32_t bounds[4];
    const std::string pre = path("im");
    EXPECT(pxl_read_image_text(pre.c_str(), W, H, 0, r.data(), g.data(), b.data(), m.data(),
                               bounds) == PXL_OK);
    EXPECT(bounds[0] == 2 && bounds[1] == 4 && bounds[2] == 3 && bounds[3] == 6);
    EXPECT(pxl_read_image_text(pre.c_str(), W, H, 1, r.data(), nullptr, nullptr, m.data(),
                               bounds) == PXL_OK);
    // One pixel more than the files hold; a token that is not a number.
    std::vector<uint8_t> r2(W * (H + 1)), g2(W * (H + 1)), b2(W * (H + 1)), m2(W * (H + 1));
    EXPECT(pxl_read_image_text(pre.c_str(), W, H + 1, 0, r2.data(), g2.data(), b2.data(),
                               m2.data(), bounds) == PXL_EINVAL);
    put("bad_mask.txt", "0 0 x 0\n");
    EXPECT(pxl_read_image_text(path("bad").c_str(), 2, 2, 1, r.data(), nullptr, nullptr, m.data(),
                               bounds) == PXL_EINVAL);
    EXPECT(pxl_read_image_text(pre.c_str(), 0, H, 0, r.data(), g.data(), b.data(), m.data(),
                               bounds) == PXL_EINVAL);
}

// A caged band-0 directory of an N^3 grid in the reference's layout.
static void caged_dir(int N, std::vector<int> &io_out, int64_t &n_out) {
    std::vector<float> dens((size_t)N * N * N, 0.0f);
    dens[(size_t)(2 * N * N + 2 * N + 2)] = 30.0f;
    dens[(size_t)(3 * N * N + 2 * N + 3)] = 25.0f;
    std::vector<uint8_t> occ((size_t)(N - 1) * (N - 1) * (N - 1)), cage((size_t)N * N * N);
    orc_build_occupancy(dens.data(), N, N, N, 10.0f, occ.data(), cage.data());
    std::vector<uint32_t> io((size_t)N * N * N);
    const int64_t n = orc_build_index(cage.data(), N, N, N, io.data());
    std::vector<int> vocc(occ.begin(), occ.end()), vcage(cage.begin(), cage.end());
    io_out.assign(io.begin(), io.end());
    put("occupancy.txt", ints(vocc));
    put("sparse_inCageInfo_band0.txt", ints(vcage));
    put("sparse_indexOffset.txt", ints(io_out));
    std::vector<float> v((size_t)n);
    for (int64_t i = 0; i < n; i++) v[(size_t)i] = 0.5f * (float)i;
    put("sparse_opacity_band0.txt", floats(v));
    for (const char *ch : {"red", "green", "blue"}) put(std::string(ch) + "_0_sparse.txt", floats(v));
    n_out = n;
}

static void grids() {
    const int N = 6;
    std::vector<int> io;
    int64_t n = 0;
    caged_dir(N, io, n);
    int64_t nrec = -1;
    EXPECT(pxl_grid_text_records(g_dir.c_str(), 0, 1, N, N, N, &nrec) == PXL_OK);
    EXPECT(pxl_grid_text_records(g_dir.c_str(), 0, 0, N, N, N, &nrec) == PXL_OK &&
           nrec == (int64_t)N * N * N);
    // The valid directory parses; the copy to the device then succeeds or,
    // with no device, reports a HIP error.  Device pointers are never
    // dereferenced on the host.
    void *dummy = reinterpret_cast<void *>(uintptr_t(4096));
    int32_t rc = pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                                    (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                                    (uint32_t *)dummy, nullptr);
    EXPECT(rc == PXL_EHIP || rc == PXL_OK);
    // More records than vertices, a record count that does not match the
    // cage, an index outside the cage, a short density file, a garbage
    // token in the index file.
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, (int64_t)N * N * N + 1,
                              (float *)dummy, (uint32_t *)dummy, (uint8_t *)dummy,
                              (uint32_t *)dummy, (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n + 1, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    std::vector<int> bad = io;
    for (size_t i = bad.size(); i-- > 0;)
        if (bad[i] == 0) {
            bad[i] = 1000000;
            break;
        }
    put("sparse_indexOffset.txt", ints(bad));
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    put("sparse_indexOffset.txt", "0 1 2 q\n");
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    put("sparse_indexOffset.txt", ints(io));
    put("sparse_opacity_band0.txt", "1.0 2.0\n");
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    put("occupancy.txt", "1 0\n");
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 0, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);
    EXPECT(pxl_grid_load_text(g_dir.c_str(), 3, 1, 1, N, N, N, 1, 4, n, (float *)dummy,
                              (uint32_t *)dummy, (uint8_t *)dummy, (uint32_t *)dummy,
                              (uint32_t *)dummy, nullptr) == PXL_EINVAL);  // band beyond sh_dim
}

static void shuffle() {
    std::vector<int32_t> d(1000);
    EXPECT(pxl_shuffle_rank(d.data(), 1000, 1, 0, 1) == PXL_OK);
    std::vector<int> seen(1000, 0);
    for (int32_t x : d)
        if (x >= 0 && x < 1000) seen[(size_t)x]++;
    for (int s : seen) EXPECT(s == 1);
}

// The oracle on a small caged grid: rays of one 12 x 12 view, forward,
// backward (plain, conditioning, masked), TV, Adam, eval image, upsampling.
static void oracle_run() {
    const int N = 12, S = 9, P = 1 + 3 * S, img = 12;
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<float> dens((size_t)N * N * N);
    for (int z = 0; z < N; z++)
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const float r2 = (float)((x - 5.5) * (x - 5.5) + (y - 6) * (y - 6) + (z - 5) * (z - 5));
                dens[(size_t)((z * N + y) * N + x)] = 40.0f * std::exp(-r2 / 10.0f);
            }
    std::vector<uint8_t> occ((size_t)(N - 1) * (N - 1) * (N - 1)), cage((size_t)N * N * N);
    orc_build_occupancy(dens.data(), N, N, N, 8.0f, occ.data(), cage.data());
    std::vector<uint32_t> io((size_t)N * N * N);
    const int64_t n = orc_build_index(cage.data(), N, N, N, io.data());
    EXPECT(n > 1);
    std::vector<float> planes((size_t)P * n);
    for (auto &v : planes) v = u(rng);
    for (int64_t r = 0; r < n; r++) planes[(size_t)r] = 10.0f + 5.0f * u(rng);
    orc_grid g{};
    g.W = g.H = g.D = N;
    g.S = S;
    g.sparse = 1;
    g.n_vox = n;
    g.planes = planes.data();
    g.occupancy = occ.data();
    g.index_offsets = io.data();
    g.in_cage = cage.data();
    const int32_t bbox[6] = {0, N - 1, 0, N - 1, 0, N - 1};
    for (int i = 0; i < 6; i++) g.bbox[i] = bbox[i];
    g.step = 0.5f;
    double c2w[16], centre[4], w2c[16];
    const double tr[3] = {2.0, 3.0, 2.5}, obj[3] = {N / 2.0, N / 2.0, N / 2.0};
    orc_camera_from_pose(tr, obj, 1.0, 0, 0, 14.0, 14.0, 6.0, 6.0, c2w, w2c, centre);
    const int R = img * img;
    std::vector<uint8_t> pix(R), msk(R, 255);
    for (int i = 0; i < R; i++) pix[(size_t)i] = (uint8_t)(i * 37);
    std::vector<float> ox(R), oy(R), oz(R), dx(R), dy(R), dz(R), tn(R), tf(R);
    std::vector<uint8_t> gr(R), gg(R), gb(R), gm(R);
    orc_rays_init(c2w, centre, 14.0, 14.0, 6.0, 6.0, bbox, img, img, pix.data(), pix.data(),
                  pix.data(), msk.data(), 0, ox.data(), oy.data(), oz.data(), dx.data(),
                  dy.data(), dz.data(), tn.data(), tf.data(), gr.data(), gg.data(), gb.data(),
                  gm.data());
    orc_rays rays{ox.data(), oy.data(), oz.data(), dx.data(), dy.data(), dz.data(), tn.data(),
                  tf.data(), gr.data(), gg.data(), gb.data(), gm.data(), 0};
    std::vector<int32_t> si(R);
    for (int i = 0; i < R; i++) si[(size_t)i] = i;
    std::vector<double> fr(R), fg(R), fb(R), fw(R), fd(R), loss(R);
    std::vector<uint32_t> trace(2 * (size_t)R);
    orc_render_fwd(&g, &rays, si.data(), 0, R, fr.data(), fg.data(), fb.data(), fw.data(),
                   fd.data(), trace.data());
    std::vector<double> grad((size_t)P * n, 0.0), gabs((size_t)P * n, 0.0);
    orc_render_bwd(&g, &rays, si.data(), 0, R, fr.data(), fg.data(), fb.data(), fw.data(), 1.0,
                   0.1f, loss.data(), grad.data());
    orc_render_bwd_cond(&g, &rays, si.data(), 0, R, fr.data(), fg.data(), fb.data(), fw.data(),
                        1.0, 0.1f, loss.data(), grad.data(), gabs.data());
    std::vector<uint64_t> bm(R), tg(R);
    for (int i = 0; i < R; i++) bm[(size_t)i] = 0x0F0F0F0F0F0F0F0Full << (i % 4);
    orc_render_fwd_masked(&g, &rays, si.data(), 0, R, bm.data(), 1e-4, tg.data(), fr.data(),
                          fg.data(), fb.data(), fw.data(), fd.data(), nullptr);
    orc_render_bwd_masked(&g, &rays, si.data(), 0, R, bm.data(), fr.data(), fg.data(), fb.data(),
                          fw.data(), 1.0, 0.1f, loss.data(), grad.data());
    orc_tv_disparity(&g, 1e-5, 1e-6, 0, grad.data());
    orc_tv_neighbour6(&g, 1e-5, 1e-6, grad.data());
    std::vector<float> p32(planes), g32((size_t)P * n), m32((size_t)P * n, 0.0f),
        v32((size_t)P * n, 0.0f);
    for (size_t i = 0; i < g32.size(); i++) g32[i] = (float)grad[i];
    orc_adam(p32.data(), g32.data(), m32.data(), v32.data(), nullptr, P, n, 0.1, 0.1, 1);
    orc_adam_band(p32.data(), g32.data(), m32.data(), v32.data(), nullptr, P, S, n, 0.1, 0.1, 2,
                  1, 4);
    std::vector<double> rgb(3 * (size_t)R), disp(R), rays_out(5 * (size_t)R);
    std::vector<uint8_t> rgb8(3 * (size_t)R), m8(R);
    orc_render_image(&g, c2w, centre, 14.0, 14.0, 6.0, 6.0, img, img, 1, rgb.data(),
                     rgb8.data(), m8.data(), disp.data(), rays_out.data());
    const int N2 = 2 * N - 1;
    std::vector<float> d2((size_t)N2 * N2 * N2);
    orc_upsample_density(&g, N2, N2, N2, 8.0f, d2.data());
    std::vector<float> up((size_t)P * N2 * N2 * N2, 0.0f);
    orc_upsample(&g, N2, N2, N2, S, nullptr, nullptr, (int64_t)N2 * N2 * N2, 8.0f, 0, up.data());
    EXPECT(std::isfinite(orc_mse(loss.data(), R)));
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: san_driver <scratch dir>\n");
        return 2;
    }
    g_dir = argv[1];
    cameras();
    images();
    grids();
    shuffle();
    oracle_run();
    if (g_fail) {
        std::fprintf(stderr, "san: %d expectation(s) failed\n", g_fail);
        return 1;
    }
    std::printf("san: ok\n");
    return 0;
}
