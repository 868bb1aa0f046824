"""Write the trajectory-file golden outputs of tests/golden/ with the
reference's own writer (ResultTxt::save, tracker/calib/ResultTxt.cpp:52-76,
compiled into oracle/_ref/libref.so by oracle/Makefile target `ref`), and copy
the reference's sample trajectory and schedule files (bin/data/
cameraMatrixEst.txt, best-method.tst). Runs in the build container only; the
outputs are data.
"""
import ctypes
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF_DATA = "/root/reference/bin/data"


def main():
    import oracle
    oracle.build(with_reference=True)
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref/libref.so is not built (needs the reference tree)")
    ref.ref_result_txt_save.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                        ctypes.c_int]
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(2018)
    n = 12
    poses = np.zeros((n, 16), np.float32)
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(rng.normal(0, 0.3, 3)).as_matrix()
    T[:3, 3] = rng.normal(0, 800, 3)
    for i in range(n):
        poses[i] = T.astype(np.float32).ravel()
        D = np.eye(4)
        D[:3, :3] = Rotation.from_rotvec(rng.normal(0, 0.02, 3)).as_matrix()
        D[:3, 3] = rng.normal(0, 15, 3)
        T = T @ D
    np.save(os.path.join(HERE, "ref_trajectory_poses.npy"), poses)
    for t in (0, 1):
        for inv in (0, 1):
            p = os.path.join(HERE, f"ref_trajectory_t{t}_i{inv}.txt")
            a = poses.copy()
            ref.ref_result_txt_save(p.encode(), n, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), t, inv)
    for name in ("cameraMatrixEst.txt", "best-method.tst"):
        shutil.copyfile(os.path.join(REF_DATA, name), os.path.join(HERE, name))
    print("wrote trajectory goldens and copied the sample files")


if __name__ == "__main__":
    main()
