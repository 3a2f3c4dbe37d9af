"""Writes tests/golden/poses_lego.npz from the reference's Poses.zip.

Runs where the reference tree is present; the tests read only the fixture:
    python tests/golden/make_poses.py /root/reference/Poses.zip
The fixture holds, per split (train / val / test) of the lego scene: the
intrinsic matrix K (cameraInfo.txt line 2-4), each view's camera translation,
its stored 4x4 world-to-camera matrix, and the camera centre listed in
centerInfo.txt.  Data only; the reference's code is not involved.
"""
import os
import sys
import zipfile

import numpy as np


def parse_camera_info(text):
    tok = text.split()
    assert tok[0] == "total"
    n = int(tok[1])
    i = 2
    K = np.array(list(map(float, tok[i:i + 9])))
    i += 9
    T, M = [], []
    for v in range(n):
        assert tok[i] == f"{v}:", (tok[i], v)
        i += 1
        T.append(list(map(float, tok[i:i + 3])))
        i += 3
        M.append(list(map(float, tok[i:i + 16])))
        i += 16
    return K, np.array(T), np.array(M).reshape(n, 4, 4)


def main(zip_path):
    z = zipfile.ZipFile(zip_path)
    out = {}
    for split in ("train", "val", "test"):
        K, T, M = parse_camera_info(z.read(f"lego/{split}/cameraInfo.txt").decode())
        C = np.array([list(map(float, ln.split()))
                      for ln in z.read(f"lego/{split}/centerInfo.txt").decode().splitlines()
                      if ln.strip()])
        assert C.shape == T.shape, (split, C.shape, T.shape)
        out[f"{split}_K"] = K
        out[f"{split}_t"] = T
        out[f"{split}_w2c"] = M
        out[f"{split}_centre"] = C
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "poses_lego.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/Poses.zip")
