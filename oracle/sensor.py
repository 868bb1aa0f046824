"""CPU restatement of the reference's sensor-side inputs: the calibration file
(calib.xml) -> calibration block, and the Kinect disparity -> depth
conversion.

TEST INFRASTRUCTURE ONLY: the checker for the library's rt_calib_* and
rt_disparity_to_depth. FP32 steps are numpy float32 operations (one IEEE
rounding each, in the reference's order); FP64 steps are Python floats.
"""
import re

import numpy as np

F = np.float32


def _elements(text):
    """name -> (size tuple, list of floats) of the file's elements."""
    out = {}
    for m in re.finditer(r"<(\w+)\b([^>]*)>([^<]*)</\1>", text):
        name, attrs, body = m.group(1), m.group(2), m.group(3)
        if name in out:
            continue
        sz = re.search(r'size="(\d+)\s+(\d+)"', attrs)
        size = (int(sz.group(1)), int(sz.group(2))) if sz else (0, 0)
        out[name] = (size, [float(v) for v in body.split()])
    return out


def rodrigues(om):
    """Rotation matrix of a rotation vector (cv::Rodrigues, used at
    calib.cpp:649-656), here through scipy's own implementation."""
    from scipy.spatial.transform import Rotation
    return Rotation.from_rotvec(np.asarray(om, np.float64)).as_matrix()


def invert_rt4(M):
    """invertRT4 (basic_math.cpp:335-353), FP64."""
    M = np.asarray(M, np.float64).ravel()
    Mi = np.zeros(16)
    Mi[0], Mi[1], Mi[2] = M[0], M[4], M[8]
    Mi[3] = -(M[0] * M[3] + M[4] * M[7] + M[8] * M[11])
    Mi[4], Mi[5], Mi[6] = M[1], M[5], M[9]
    Mi[7] = -(M[1] * M[3] + M[5] * M[7] + M[9] * M[11])
    Mi[8], Mi[9], Mi[10] = M[2], M[6], M[10]
    Mi[11] = -(M[2] * M[3] + M[6] * M[7] + M[10] * M[11])
    Mi[15] = 1.0
    return Mi


def inverse3x3_f32(m):
    """det3x3 + inverse3x3 for floats (basic_math.cpp:507-514, 562-582): the
    identity when |det| < 1e-5."""
    m = [F(v) for v in np.asarray(m, np.float32).ravel()]
    det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[1] * (m[3] * m[8] - m[6] * m[5]) + m[2] * (m[3] * m[7] - m[6] * m[4])
    if abs(det) < F(1e-5):
        return np.eye(3, dtype=np.float32).ravel()
    t = [(m[4] * m[8] - m[5] * m[7]) / det, -(m[1] * m[8] - m[7] * m[2]) / det, (m[1] * m[5] - m[4] * m[2]) / det,
         -(m[3] * m[8] - m[5] * m[6]) / det, (m[0] * m[8] - m[6] * m[2]) / det, -(m[0] * m[5] - m[3] * m[2]) / det,
         (m[3] * m[7] - m[6] * m[4]) / det, -(m[0] * m[7] - m[6] * m[1]) / det, (m[0] * m[4] - m[1] * m[3]) / det]
    return np.array(t, np.float32)


def calib_block(text, width, height):
    """Calibration::init (calib.cpp:521-716) + setupCalibDataBuffer(width,
    height) (calib.cpp:73-93) of a calib.xml text: (block float32[96],
    xy_offset)."""
    el = _elements(text)
    for req in ("KL", "KR", "om", "T"):
        if req not in el:
            raise ValueError(f"{req} not found")
    KL = np.array(el["KL"][1][:9]).reshape(3, 3).T.ravel()  # stored column by column (calib.cpp:581-582)
    KR = np.array(el["KR"][1][:9]).reshape(3, 3).T.ravel()
    kcR = np.zeros(8)
    if "kc_right" in el:
        v = el["kc_right"][1]
        kcR[:len(v)] = v
    B = el["B"][1][0] if "B" in el else 0.0
    b = el["b"][1][0] if "b" in el else 0.0
    TLR = np.eye(4)
    TLR[:3, :3] = rodrigues(el["om"][1][:3])
    TLR[:3, 3] = el["T"][1][:3]
    TLR = TLR.ravel()
    TRL = invert_rt4(TLR)
    scale = 1.0 / (8 * b * abs(KL[0]) * 640.0)  # calib.cpp:699-701 (double)
    c1, c0 = -scale, B * 640.0 * scale
    c = np.zeros(96, np.float32)

    def K_block(K):
        k = np.eye(3, dtype=np.float32).ravel()
        for i in range(6):
            k[i] = F(width * K[i])
        return k, inverse3x3_f32(k)
    c[0:9], c[9:18] = K_block(KR)
    c[18:26] = kcR.astype(np.float32)
    c[26:42] = TLR.astype(np.float32)
    c[64:80] = TRL.astype(np.float32)
    c[43:52], c[52:61] = K_block(KL)
    c[61], c[42] = F(c0), F(c1)
    c[62], c[63] = F(300.0), F(7000.0)
    if "zpoly" in el:
        z = el["zpoly"][1]
        c[80:80 + len(z)] = np.array(z, np.float32)
    return c, 1


def disparity_to_depth(disp, calib, xy_offset=True, half=True, mode=0):
    """OMPFunctions::d2ZLow / d2Z / d2ZLowGPU (multicore.cpp:189-379): depth in
    mm (mode 0) or normalised to the depth window (mode 1); 0 where invalid;
    rows from 4*floor(out_h/4) on are 0 (never written by the reference's four
    row blocks). Source coordinates below 0 are outside (the CPU conversion)."""
    d = np.asarray(disp, np.uint16)
    h, w = d.shape
    oh, ow = (h // 2, w // 2) if half else (h, w)
    ox, oy = (-4, -3) if xy_offset else (0, 0)
    c0, c1 = F(calib[61]), F(calib[42])
    mn, mx = F(calib[62]), F(calib[63])
    out = np.zeros((oh, ow), np.float32)
    for y in range(4 * (oh // 4)):
        for x in range(ow):
            sx = (2 * x if half else x) + ox
            sy = (2 * y if half else y) + oy
            if not (0 <= sx < w - 1 and 0 <= sy < h - 1):
                continue
            if half:
                vals = [F(d[sy, sx]), F(d[sy, sx + 1]), F(d[sy + 1, sx + 1]), F(d[sy + 1, sx])]
            else:
                vals = [F(d[sy, sx])]
            if not all(v < F(2047) for v in vals):
                continue
            dd = vals[0]
            for v in vals[1:]:
                if v < dd:
                    dd = v
            z = abs(F(1.0) / (c0 + c1 * dd))
            if z > mx or z < mn:
                z = F(0.0)
            out[y, x] = (z - mn) / (mx - mn) if mode else z
    return out


def disparity_to_depth_np(disp, calib, xy_offset=True, half=True, mode=0):
    """The same, vectorised (the same float32 operations element by element)."""
    d = np.asarray(disp, np.uint16)
    h, w = d.shape
    oh, ow = (h // 2, w // 2) if half else (h, w)
    ox, oy = (-4, -3) if xy_offset else (0, 0)
    c0, c1 = F(calib[61]), F(calib[42])
    mn, mx = F(calib[62]), F(calib[63])
    ys, xs = np.mgrid[0:oh, 0:ow]
    sx = (2 * xs if half else xs) + ox
    sy = (2 * ys if half else ys) + oy
    inside = (sx >= 0) & (sy >= 0) & (sx < w - 1) & (sy < h - 1) & (ys < 4 * (oh // 4))
    sxc, syc = np.clip(sx, 0, w - 2), np.clip(sy, 0, h - 2)
    d1 = d[syc, sxc].astype(np.float32)
    if half:
        d2 = d[syc, sxc + 1].astype(np.float32)
        d3 = d[syc + 1, sxc + 1].astype(np.float32)
        d4 = d[syc + 1, sxc].astype(np.float32)
        ok = (d1 < 2047) & (d2 < 2047) & (d3 < 2047) & (d4 < 2047)
        dd = np.minimum(np.minimum(np.minimum(d1, d2), d3), d4)
    else:
        ok = d1 < 2047
        dd = d1
    z = np.abs(F(1.0) / (c0 + c1 * dd)).astype(np.float32)
    z = np.where((z > mx) | (z < mn), F(0.0), z).astype(np.float32)
    v = ((z - mn) / (mx - mn)).astype(np.float32) if mode else z
    return np.where(inside & ok, v, F(0.0)).astype(np.float32)
