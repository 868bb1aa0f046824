"""Sensor-side inputs in the reference's formats through the C ABI
(include/rt_sensor.h): calibration files (calib.xml), PNM frame files, and the
Kinect disparity -> depth conversion on the GPU."""
import ctypes

import numpy as np

from . import lib
from .hip import DeviceBuffer, hip

_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


def _check(rc, what):
    if rc != 0:
        msg = lib().rt_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def load_calib_xml(path, width, height):
    """(calibration block float32[96], xy_offset flag) of a calib.xml file at
    the given image size."""
    c = np.zeros(96, np.float32)
    off = ctypes.c_int(0)
    _check(lib().rt_calib_load_xml(str(path).encode(), width, height, c.ctypes.data_as(_fp), ctypes.byref(off)),
           "rt_calib_load_xml")
    return c, off.value


def parse_calib_xml(text, width, height):
    """The same from the file's text."""
    b = text.encode() if isinstance(text, str) else bytes(text)
    c = np.zeros(96, np.float32)
    off = ctypes.c_int(0)
    _check(lib().rt_calib_parse_xml(b, len(b), width, height, c.ctypes.data_as(_fp), ctypes.byref(off)),
           "rt_calib_parse_xml")
    return c, off.value


def read_pnm(path):
    """Samples of a binary P5/P6 file: uint8 or uint16 (host order), shape
    (h, w) or (h, w, 3)."""
    w, h, ch, mx = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = str(path).encode()
    _check(lib().rt_pnm_info(p, ctypes.byref(w), ctypes.byref(h), ctypes.byref(ch), ctypes.byref(mx)), "rt_pnm_info")
    a = np.empty((h.value, w.value, ch.value), np.uint16 if mx.value > 255 else np.uint8)
    _check(lib().rt_pnm_read(p, a.ctypes.data, a.nbytes), "rt_pnm_read")
    return a[:, :, 0] if ch.value == 1 else a


def disparity_to_depth(disparity, calib, xy_offset=True, half=True, mode=0, out=None, stream=None, sync=True):
    """Depth map (float32, mm for mode 0, normalised for mode 1) of a Kinect
    disparity frame. disparity: host uint16 array (h, w), or a tuple
    (DeviceBuffer, h, w) of device data; returns a host array, or fills `out`
    (a DeviceBuffer) and returns it."""
    if isinstance(disparity, tuple):
        d, h, w = disparity
    else:
        a = np.ascontiguousarray(disparity, np.uint16)
        h, w = a.shape
        d = DeviceBuffer.from_array(a)
    oh, ow = (h // 2, w // 2) if half else (h, w)
    o = out if out is not None else DeviceBuffer(ow * oh * 4)
    c = np.ascontiguousarray(calib, np.float32)
    _check(lib().rt_disparity_to_depth(d.ptr, w, h, c.ctypes.data_as(_fp), int(bool(xy_offset)), int(bool(half)),
                                       mode, o.ptr, stream), "rt_disparity_to_depth")
    if out is not None:
        return out
    if sync:
        hip().hipDeviceSynchronize()
    return o.to_array(np.float32, ow * oh).reshape(oh, ow)


def depth_mm_to_u16(depth_dev, n, out_dev, stream=None):
    """Float mm -> the tracker's 16-bit depth input (device buffers)."""
    _check(lib().rt_depth_mm_to_u16(depth_dev.ptr, n, out_dev.ptr, stream), "rt_depth_mm_to_u16")


def bayer_to_gray_half(bayer):
    """8-bit gray at half resolution from a Kinect Bayer mosaic (the layout of
    the reference's bayer_rgbimageNNNN.ppm files, OpenCV's BayerGB: G R / B G):
    each 2x2 cell gives round(0.3 R + 0.59 (G1 + G2) / 2 + 0.11 B), the gray
    weights of rgbTex2GrayHdrZKernel (rgb2gray.cu:64-76). A simplified stand-in
    for the reference's capture path (OpenCV demosaic, pyrDown, then the gray
    conversion), which this library does not reproduce."""
    b = np.asarray(bayer, np.float64)
    g1, r, bl, g2 = b[0::2, 0::2], b[0::2, 1::2], b[1::2, 0::2], b[1::2, 1::2]
    return np.clip(np.rint(0.3 * r + 0.59 * 0.5 * (g1 + g2) + 0.11 * bl), 0, 255).astype(np.uint8)
