"""Host-side text formats (no GPU): camera files (Camera::readFromMyBlender,
SparseVolumeRadianceField/CameraUtils.cu:271-386) and image text
(Image::initSelfFromFile + readMask, ImageUtils.cu:84-128, Utils.cu:508-572,
written by utils/preprocess_datasets.py:9-22 with np.savetxt fmt '%d').

Pins: the lego camera files of the committed pose fixture (Poses.zip)
reproduce the stored camera centres and rotations; the intrinsics rescale is
the closed form K <- diag(target/2c) K; image planes and mask bounds against
numpy on files in the preprocessing script's layout."""
import os

import numpy as np
import pytest

import pxl

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _write_camera_file(path, K, T, W2C):
    """cameraInfo.txt layout: 'total N', the 3x3 intrinsics, then per view
    'i:', the translation (one number a line) and the 4x4 matrix."""
    with open(path, "w") as f:
        f.write(f"total {len(T)}\n")
        for r in range(3):
            f.write(" ".join(f"{K[r * 3 + c]:.6f}" for c in range(3)) + "\n")
        for i in range(len(T)):
            f.write(f"{i}:\n")
            for v in T[i]:
                f.write(f"{v:.6f}\n")
            for r in range(4):
                f.write(" ".join(f"{W2C[i][r, c]:.6f}" for c in range(4)) + "\n")


@pytest.mark.parametrize("split", ["train", "test"])
def test_read_cameras_lego(tmp_path, split):
    d = np.load(os.path.join(GOLDEN, "poses_lego.npz"))
    K, T, W2C, C = d[f"{split}_K"], d[f"{split}_t"], d[f"{split}_w2c"], d[f"{split}_centre"]
    path = tmp_path / "cameraInfo.txt"
    _write_camera_file(path, K, T, W2C)
    Kf = np.array([float(f"{k:.6f}") for k in K])
    cams = pxl.read_cameras(path, (0.0, 0.0, 0.0), 800, 800, 4.031129, need_rescale=False,
                            z_reversed=False)
    assert len(cams) == len(T)
    for i, cam in enumerate(cams):
        t = [float(f"{v:.6f}") for v in T[i]]
        ref, _ = pxl.camera_from_pose(t, (0.0, 0.0, 0.0), 4.031129, Kf[0], Kf[4], Kf[2], Kf[5],
                                      need_rescale=False, z_reversed=False)
        assert bytes(cam) == bytes(ref)  # the file reader adds nothing to the pose path
        assert np.abs(np.array(cam.centre[:3]) - C[i]).max() < 1e-5
        assert (cam.fx, cam.fy, cam.cx, cam.cy) == (Kf[0], Kf[4], Kf[2], Kf[5])
    # camera_nums limits the views read
    assert len(pxl.read_cameras(path, (0.0, 0.0, 0.0), 800, 800, 4.031129, camera_nums=7)) == 7


def test_read_cameras_rescaled_intrinsics(tmp_path):
    d = np.load(os.path.join(GOLDEN, "poses_lego.npz"))
    K, T, W2C = d["val_K"], d["val_t"][:3], d["val_w2c"][:3]
    path = tmp_path / "cameraInfo.txt"
    _write_camera_file(path, K, T, W2C)
    Kf = np.array([float(f"{k:.6f}") for k in K])
    # principal point 400 against a 200 x 100 target: sx = 200/400/2, sy = 100/400/2
    cams = pxl.read_cameras(path, (128.0,) * 3, 200, 100, 4.031129)
    sx, sy = 200 / Kf[2] / 2, 100 / Kf[5] / 2
    for cam in cams:
        assert cam.fx == sx * Kf[0] and cam.cx == sx * Kf[2] == 100.0
        assert cam.fy == sy * Kf[4] and cam.cy == sy * Kf[5] == 50.0


def test_read_cameras_rejects_corrupt(tmp_path):
    d = np.load(os.path.join(GOLDEN, "poses_lego.npz"))
    path = tmp_path / "cameraInfo.txt"
    _write_camera_file(path, d["val_K"], d["val_t"][:3], d["val_w2c"][:3])
    text = path.read_text().replace("1:\n", "2:\n", 1)
    path.write_text(text)
    with pytest.raises(pxl.PxlError):
        pxl.read_cameras(path, (0.0,) * 3, 800, 800, 4.031129)


def test_read_image_text(tmp_path):
    rng = np.random.default_rng(2)
    H, W = 30, 40
    planes = {c: rng.integers(0, 256, (H, W)).astype(np.uint8) for c in ("red", "green", "blue")}
    mask = np.zeros((H, W), np.uint8)
    mask[7:21, 3:36] = rng.integers(0, 256, (14, 33))
    mask[7, 10] = mask[20, 20] = mask[12, 3] = mask[15, 35] = 255
    prefix = tmp_path / "r_0"
    np.savetxt(f"{prefix}_mask.txt", mask, fmt="%d")
    for c, v in planes.items():
        np.savetxt(f"{prefix}_{c}.txt", v, fmt="%d")
    gray = rng.integers(0, 256, (H, W)).astype(np.uint8)
    np.savetxt(f"{prefix}.txt", gray, fmt="%d")
    out = pxl.read_image_text(prefix, W, H)
    assert np.array_equal(out["mask"], mask.reshape(-1))
    assert np.array_equal(out["r"], planes["red"].reshape(-1))
    assert np.array_equal(out["g"], planes["green"].reshape(-1))
    assert np.array_equal(out["b"], planes["blue"].reshape(-1))
    rows, cols = np.nonzero(mask)
    assert out["bounds"] == [rows.min(), rows.max(), cols.min(), cols.max()]
    g = pxl.read_image_text(prefix, W, H, gray_only=True)
    assert np.array_equal(g["r"], gray.reshape(-1))
    # an all-zero mask keeps the start values (height - 1, 0, width - 1, 0)
    np.savetxt(f"{prefix}_mask.txt", np.zeros((H, W), np.uint8), fmt="%d")
    assert pxl.read_image_text(prefix, W, H)["bounds"] == [H - 1, 0, W - 1, 0]
    with pytest.raises(pxl.PxlError):
        pxl.read_image_text(prefix, W, H + 1)


def _write_caged_dir(d, occ, cage, io, n):
    """A band-0 caged directory in the reference's file layout
    (objectScene::saveSelf, SceneUtils.cu:1246-1299)."""
    def ints(name, a):
        (d / name).write_text(" ".join(str(int(x)) for x in a) + "\n")

    def floats(name, a):
        (d / name).write_text(" ".join(f"{x:.6f}" for x in a) + "\n")

    ints("occupancy.txt", occ)
    ints("sparse_inCageInfo_band0.txt", cage)
    ints("sparse_indexOffset.txt", io)
    rng = np.random.default_rng(3)
    floats("sparse_opacity_band0.txt", rng.uniform(0, 20, n))
    for ch in ("red", "green", "blue"):
        floats(f"{ch}_0_sparse.txt", rng.normal(0, 1, n))


@pytest.mark.parametrize("corrupt", ["outside", "first"])
def test_grid_load_text_rejects_bad_offsets(tmp_path, corrupt):
    """A caged directory whose index file sends a vertex outside the cage
    (or the first in-cage vertex) to a record other than 0 is rejected
    before anything reaches the device: kernels would otherwise read and
    write records that do not exist (getCagedNums gives those vertices 0,
    SceneUtils.cu:1106-1113)."""
    import ctypes

    import oracle

    N = 5
    dens = np.zeros(N ** 3, np.float32)
    dens[2 * N * N + 2 * N + 2] = 30.0
    occ, cage, _ = oracle.build_occupancy(dens, N, N, N, 10.0)
    io, n = oracle.build_index(cage, N, N, N)
    io = io.astype(np.int64)
    if corrupt == "outside":
        io[int(np.flatnonzero(cage == 0)[-1])] = 10 ** 6
    else:
        io[int(np.flatnonzero(cage)[0])] = 1
    _write_caged_dir(tmp_path, occ, cage, io, n)
    L = pxl.lib()
    dummy = ctypes.c_void_p(16)  # never reached: validation fails first
    rc = L.pxl_grid_load_text(str(tmp_path).encode(), 0, 1, 1, N, N, N, 1, 4, n, dummy, dummy,
                              dummy, dummy, dummy, None)
    assert rc == pxl.PXL_EINVAL
    msg = L.pxl_last_error()
    assert (b"outside the cage" in msg) if corrupt == "outside" else (b"first in-cage" in msg), msg
