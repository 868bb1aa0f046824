"""Write tests/golden/kinect_frames.npz and tests/golden/calib.xml from the
reference's sample data (bin/data/sequence/bayer_rgbimageNNNN.ppm,
rawdepthNNNN.ppm and bin/data/calib.xml): the raw 8-bit Bayer mosaics and
11-bit Kinect disparity frames 1-4, unchanged, and the calibration file. Runs
in the build container only (the reference tree is not present where the GPU tests run); the
fixture holds data only.
"""
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/bin/data"


def read_pnm(path):
    b = open(path, "rb").read()
    fields, i = [], 2
    while len(fields) < 3:
        while b[i:i + 1].isspace():
            i += 1
        j = i
        while not b[j:j + 1].isspace():
            j += 1
        fields.append(int(b[i:j]))
        i = j
    i += 1
    w, h, mx = fields
    return np.frombuffer(b[i:], ">u2" if mx > 255 else "u1").reshape(h, w).astype(np.uint16 if mx > 255 else np.uint8)


def main():
    if not os.path.isdir(REF):
        sys.exit("the reference's sample data is not present")
    out = {}
    for k in range(1, 5):
        out[f"bayer{k}"] = read_pnm(os.path.join(REF, "sequence", f"bayer_rgbimage{k:04d}.ppm"))
        out[f"disparity{k}"] = read_pnm(os.path.join(REF, "sequence", f"rawdepth{k:04d}.ppm"))
    np.savez_compressed(os.path.join(HERE, "kinect_frames.npz"), **out)
    shutil.copyfile(os.path.join(REF, "calib.xml"), os.path.join(HERE, "calib.xml"))
    print({k: (v.shape, v.dtype) for k, v in out.items()})


if __name__ == "__main__":
    main()
