// Host-side camera construction: the part of Camera (CameraUtils.cu) that
// ray generation consumes - look-at rotation, OpenGL->OpenCV flip,
// world-to-camera / camera-to-world matrices and the camera centre.
#include <cmath>
#include <cstring>

#include "pxl.h"
#include "pxl_host.hpp"

namespace pxl {
namespace host {

namespace {

constexpr double kPi = 3.141592653589793;

// out[n x m] = a[n x k] * b[k x m].
void matmul(const double *a, const double *b, double *out, int n, int k, int m) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < m; j++) {
            double s = 0.0;
            for (int t = 0; t < k; t++) s += a[i * k + t] * b[t * m + j];
            out[i * m + j] = s;
        }
}

}  // namespace

// Gauss-Jordan elimination on [A | I]; a pivot smaller than 1e-8 is
// replaced by the next rows in turn, as getMatInvert does (Utils.cu:233-299).
bool invert(const double *src, double *dst, int n) {
    double a[4][8];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 2 * n; j++) a[i][j] = j < n ? src[i * n + j] : (j - n == i ? 1.0 : 0.0);
    for (int r = 0; r < n; r++) {
        int tries = 0;
        while (std::fabs(a[r][r]) < 1e-8) {
            ++tries;
            if (r + tries >= n) return false;
            for (int j = 0; j < 2 * n; j++) std::swap(a[r][j], a[r + tries][j]);
        }
        const double piv = a[r][r];
        for (int j = r; j < 2 * n; j++) a[r][j] /= piv;
        for (int s = r + 1; s < n; s++) {
            const double f = a[s][r];
            for (int j = 0; j < 2 * n; j++) a[s][j] -= f * a[r][j];
        }
    }
    for (int r = n - 2; r >= 0; r--)
        for (int s = n - 1; s > r; s--) {
            const double f = a[r][s];
            for (int j = r + 1; j < 2 * n; j++) a[r][j] -= f * a[s][j];
        }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) dst[i * n + j] = a[i][n + j];
    return true;
}

// Rotation from Euler angles in degrees: inverse of Rx Ry Rz
// (eulerAngle2RotationMat, Utils.cu:182-231).
void euler_to_rotation(const double deg[3], double *R) {
    const double ax = deg[0] * kPi / 180.0, ay = deg[1] * kPi / 180.0, az = deg[2] * kPi / 180.0;
    const double rx[9] = {1, 0, 0, 0, std::cos(ax), -std::sin(ax), 0, std::sin(ax), std::cos(ax)};
    const double ry[9] = {std::cos(ay), 0, std::sin(ay), 0, 1, 0, -std::sin(ay), 0, std::cos(ay)};
    const double rz[9] = {std::cos(az), -std::sin(az), 0, std::sin(az), std::cos(az), 0, 0, 0, 1};
    double rxy[9], m[9];
    matmul(rx, ry, rxy, 3, 3, 3);
    matmul(rxy, rz, m, 3, 3, 3);
    invert(m, R, 3);
}

namespace {

void unit(double *v) {
    const double s = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] /= s;
    v[1] /= s;
    v[2] /= s;
}

void cross(const double *a, const double *b, double *o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Camera level with its target (setCameraLookAtXOY, Utils.cu:404-446).
void look_at_level(const double *cam, const double *obj, double *R) {
    double z[3] = {cam[0] - obj[0], cam[1] - obj[1], cam[2] - obj[2]};
    if (std::fabs(z[0]) < 1e-8 && std::fabs(z[2]) < 1e-8) {
        const double ang[3] = {z[1] < 0.0 ? 90.0 : -90.0, 0.0, z[1] < 0.0 ? 0.0 : 180.0};
        euler_to_rotation(ang, R);
        return;
    }
    unit(z);
    const double up[3] = {0.0, 1.0, 0.0};
    double x[3], y[3];
    cross(up, z, x);
    unit(x);
    cross(z, x, y);
    const double w2c[9] = {x[0], x[1], x[2], y[0], y[1], y[2], z[0], z[1], z[2]};
    double c2w[9], prev[9], rot[9];
    invert(w2c, c2w, 3);
    std::memcpy(prev, c2w, sizeof(prev));
    const double quarter[3] = {0.0, 0.0, 90.0};
    euler_to_rotation(quarter, rot);
    matmul(prev, rot, c2w, 3, 3, 3);
    const double ey[3] = {0.0, 1.0, 0.0};
    double yw[3];
    matmul(c2w, ey, yw, 3, 3, 1);
    if (yw[2] < 0.0) {
        std::memcpy(prev, c2w, sizeof(prev));
        const double half[3] = {0.0, 0.0, 180.0};
        euler_to_rotation(half, rot);
        matmul(prev, rot, c2w, 3, 3, 3);
    }
    invert(c2w, R, 3);
}

}  // namespace

// World-to-camera rotation looking from `cam` at `obj` with +z up
// (setCameraLookAt, Utils.cu:448-486).  Rows are the camera x, y, z axes;
// camera z points from the target to the camera.
void look_at(const double *cam, const double *obj, double *R) {
    const double A = cam[0] - obj[0], B = cam[1] - obj[1], C = cam[2] - obj[2];
    if (std::fabs(C) < 1e-8) {
        look_at_level(cam, obj, R);
        return;
    }
    const double bc = std::sqrt(B * B + C * C);
    const double abc = std::sqrt(A * A + B * B + C * C);
    const double s_coef = -B / bc;
    const double c_coef = -A * C / abc / bc;
    const double ang = kPi / 2 - std::atan2(c_coef, s_coef);
    double y[3] = {std::sqrt((B * B + C * C) / (A * A + B * B + C * C)) * std::cos(ang),
                   1.0 / bc * (C * std::sin(ang) - A * B / abc * std::cos(ang)),
                   s_coef * std::sin(ang) + c_coef * std::cos(ang)};
    double z[3] = {A, B, C}, x[3];
    unit(y);
    unit(z);
    cross(y, z, x);
    for (int i = 0; i < 3; i++) {
        R[i] = x[i];
        R[3 + i] = y[i];
        R[6 + i] = z[i];
    }
}

}  // namespace host
}  // namespace pxl

// Camera of one pose (readFromMyBlender per view, CameraUtils.cu:339-377,
// formulateProjectionMat :45-106, transferSelf2GPU :577-608).
extern "C" int32_t pxl_camera_from_pose(const double translation[3], const double object_centre[3],
                                        double camera_distance, int32_t need_rescale,
                                        int32_t z_reversed, double fx, double fy, double cx,
                                        double cy, pxl_camera *out, double w2c_out[16]) {
    using namespace pxl::host;
    if (!translation || !object_centre || !out || camera_distance == 0.0) return PXL_EINVAL;
    double c[3] = {translation[0], translation[1], translation[2]};
    const double radius = (double)sqrtf((float)(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]));
    if (need_rescale && radius < 10.0)
        for (int i = 0; i < 3; i++) c[i] = c[i] * (object_centre[i] / camera_distance) + object_centre[i];
    double R[9];
    look_at(c, object_centre, R);
    if (z_reversed) {
        const double flip_deg[3] = {180.0, 0.0, 0.0};
        double flip[9], fr[9];
        euler_to_rotation(flip_deg, flip);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double s = 0.0;
                for (int t = 0; t < 3; t++) s += flip[i * 3 + t] * R[t * 3 + j];
                fr[i * 3 + j] = s;
            }
        std::memcpy(R, fr, sizeof(R));
    }
    // w2c = [R | 0; 0 1] * [I | -c; 0 1]; c2w = inverse.
    const double r4[16] = {R[0], R[1], R[2], 0, R[3], R[4], R[5], 0, R[6], R[7], R[8], 0, 0, 0, 0, 1};
    const double tr[16] = {1, 0, 0, -c[0], 0, 1, 0, -c[1], 0, 0, 1, -c[2], 0, 0, 0, 1};
    double w2c[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int t = 0; t < 4; t++) s += r4[i * 4 + t] * tr[t * 4 + j];
            w2c[i * 4 + j] = s;
        }
    if (!invert(w2c, out->c2w, 4)) return PXL_EINVAL;
    // P = K [I | 0] w2c; centre = -M^-1 p4.
    const double k34[12] = {fx, 0.0, cx, 0.0, 0.0, fy, cy, 0.0, 0.0, 0.0, 1.0, 0.0};
    double P[12];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int t = 0; t < 4; t++) s += k34[i * 4 + t] * w2c[t * 4 + j];
            P[i * 4 + j] = s;
        }
    const double M[9] = {P[0], P[1], P[2], P[4], P[5], P[6], P[8], P[9], P[10]};
    double Mi[9];
    if (!invert(M, Mi, 3)) return PXL_EINVAL;
    for (int i = 0; i < 3; i++)
        out->centre[i] = -(Mi[i * 3 + 0] * P[3] + Mi[i * 3 + 1] * P[7] + Mi[i * 3 + 2] * P[11]);
    out->centre[3] = 1.0;
    out->fx = fx;
    out->fy = fy;
    out->cx = cx;
    out->cy = cy;
    if (w2c_out) std::memcpy(w2c_out, w2c, sizeof(w2c));
    return PXL_OK;
}
