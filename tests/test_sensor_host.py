"""The sensor-side inputs on the CPU: the calibration file -> calibration
block (library against the oracle, bit for bit, and the oracle pinned by the
reference's own inverse3x3, scipy's Rodrigues and the disparity model's closed
form), PNM files, and the oracle's disparity -> depth conversion against closed
forms."""
import os

import numpy as np
import pytest

from oracle import sensor as osn
from recon_amd import sensor

HERE = os.path.dirname(os.path.abspath(__file__))
CALIB_XML = os.path.join(HERE, "golden", "calib.xml")
GOLD = np.load(os.path.join(HERE, "golden", "ref_pose_math.npz"))


@pytest.mark.parametrize("width,height", [(320, 240), (640, 480)])
def test_calib_xml_matches_oracle_bitexact(width, height):
    c, off = sensor.load_calib_xml(CALIB_XML, width, height)
    o, ooff = osn.calib_block(open(CALIB_XML).read(), width, height)
    assert off == ooff == 1
    assert np.array_equal(c, o), np.flatnonzero(c != o)
    # the same through the in-memory entry point
    c2, _ = sensor.parse_calib_xml(open(CALIB_XML).read(), width, height)
    assert np.array_equal(c, c2)


def test_calib_block_values_follow_the_file():
    """Closed forms from the file's numbers: intrinsics rows scaled by the
    width, the column-major storage, row 2 = (0, 0, 1), TLR a rotation with the
    file's translation and TRL its inverse."""
    el = osn._elements(open(CALIB_XML).read())
    c, _ = sensor.load_calib_xml(CALIB_XML, 320, 240)
    KR = np.array(el["KR"][1]).reshape(3, 3).T
    assert np.allclose(c[0:9].reshape(3, 3)[:2], (320 * KR)[:2], rtol=1e-7)
    assert list(c[6:9]) == [0, 0, 1]
    assert c[0] < 0 and c[4] > 0  # OpenGL convention of the file
    TLR, TRL = c[26:42].reshape(4, 4).astype(np.float64), c[64:80].reshape(4, 4).astype(np.float64)
    assert np.allclose(TLR[:3, :3] @ TLR[:3, :3].T, np.eye(3), atol=1e-6)
    assert np.allclose(TLR[:3, 3], el["T"][1], rtol=1e-6)
    assert np.allclose(TLR @ TRL, np.eye(4), atol=1e-4)
    assert np.allclose(c[18:23], el["kc_right"][1], rtol=1e-7)


def test_oracle_inverse3x3_equals_the_reference_bitexact():
    K, Ki = GOLD["inv3_K"], GOLD["inv3_Ki"]
    for k, ki in zip(K, Ki):
        assert np.array_equal(osn.inverse3x3_f32(k), ki)
    # and is the inverse where the determinant is not tiny
    for k, ki in zip(K[:20], Ki[:20]):
        assert np.allclose(ki.reshape(3, 3) @ k.reshape(3, 3), np.eye(3), atol=2e-4)
    # the identity branch is taken for the nearly singular ones
    assert any(np.array_equal(ki, np.eye(3, dtype=np.float32).ravel()) for ki in Ki[20:])


def test_oracle_rodrigues_closed_form():
    for ax, t in [((0, 0, 1), 0.3), ((1, 0, 0), -1.2), ((0, 1, 0), 2.5)]:
        R = osn.rodrigues(np.array(ax) * t)
        c, s = np.cos(t), np.sin(t)
        E = {(0, 0, 1): [[c, -s, 0], [s, c, 0], [0, 0, 1]], (1, 0, 0): [[1, 0, 0], [0, c, -s], [0, s, c]],
             (0, 1, 0): [[c, 0, s], [0, 1, 0], [-s, 0, c]]}[ax]
        assert np.allclose(R, E, atol=1e-15)
    assert np.array_equal(osn.rodrigues([0, 0, 0]), np.eye(3))


def test_disparity_constants_follow_the_kinect_model():
    """c0 + c1 d = (640 B - d) / (8 b |fx| 640): z = 8 b |fx'| / (B' - d), the
    commented-out form of d2z.cu:153 with fx' = 640 |KL[0]|, B' = 640 B."""
    el = osn._elements(open(CALIB_XML).read())
    B, b, fx = el["B"][1][0], el["b"][1][0], abs(el["KL"][1][0])
    c, _ = sensor.load_calib_xml(CALIB_XML, 320, 240)
    for d in (400.0, 700.0, 1000.0):
        z = 1.0 / (np.float64(c[61]) + np.float64(c[42]) * d)
        assert abs(z - 8 * b * fx * 640.0 / (B * 640.0 - d)) < 1e-6 * z


def test_calib_xml_errors():
    text = open(CALIB_XML).read()
    with pytest.raises(RuntimeError):
        sensor.parse_calib_xml(text.replace("<KL ", "<KX ").replace("</KL>", "</KX>"), 320, 240)
    with pytest.raises(RuntimeError):
        sensor.parse_calib_xml(text.replace("<om ", "<ox ").replace("</om>", "</ox>"), 320, 240)
    # kc_right is optional (zeros), zpoly lands at entries 80..
    t2 = text.replace("<kc_right", "<kc_unused").replace("</kc_right>", "</kc_unused>")
    t2 = t2.replace("</root>", '<zpoly idx="1" type="double" size="3 1">1.5 -2.0 0.25</zpoly>\n</root>')
    c, _ = sensor.parse_calib_xml(t2, 320, 240)
    o, _ = osn.calib_block(t2, 320, 240)
    assert np.array_equal(c, o)
    assert list(c[18:26]) == [0] * 8 and list(c[80:83]) == [1.5, -2.0, 0.25]


def _write_pnm(path, a, magic, maxval, comment=False):
    h, w = a.shape[:2]
    hdr = f"{magic}\n" + ("# written by the test\n" if comment else "") + f"{w} {h}\n{maxval}\n"
    data = a.astype(">u2").tobytes() if maxval > 255 else a.astype(np.uint8).tobytes()
    open(path, "wb").write(hdr.encode() + data)


def test_pnm_read(tmp_path):
    rng = np.random.default_rng(3)
    g8 = rng.integers(0, 256, (7, 11), dtype=np.uint8)
    g16 = rng.integers(0, 2048, (5, 9), dtype=np.uint16)
    rgb = rng.integers(0, 256, (4, 6, 3), dtype=np.uint8)
    _write_pnm(tmp_path / "a.pgm", g8, "P5", 255, comment=True)
    _write_pnm(tmp_path / "b.pgm", g16, "P5", 65535)
    _write_pnm(tmp_path / "c.ppm", rgb, "P6", 255)
    assert np.array_equal(sensor.read_pnm(tmp_path / "a.pgm"), g8)
    assert np.array_equal(sensor.read_pnm(tmp_path / "b.pgm"), g16)
    assert np.array_equal(sensor.read_pnm(tmp_path / "c.ppm"), rgb)
    (tmp_path / "bad.pgm").write_bytes(b"P2\n1 1\n255\n0")
    with pytest.raises(RuntimeError):
        sensor.read_pnm(tmp_path / "bad.pgm")


def test_oracle_disparity_to_depth_closed_forms():
    c, _ = sensor.load_calib_xml(CALIB_XML, 320, 240)
    c0, c1, mn, mx = np.float32(c[61]), np.float32(c[42]), np.float32(c[62]), np.float32(c[63])
    d = np.full((48, 64), 800, np.uint16)
    z = osn.disparity_to_depth(d, c, xy_offset=False, half=True)
    expect = np.abs(np.float32(1) / (c0 + c1 * np.float32(800)))
    assert z.shape == (24, 32)
    assert np.all(z == expect)  # sources (2x, 2y) + {0,1}^2 all inside: 2*31 + 1 = 63 < 64
    # an invalid disparity anywhere in a 2x2 cell empties it; otherwise the cell takes the smallest disparity
    d2 = d.copy(); d2[10, 10] = 2047; d2[20, 20] = 600
    z2 = osn.disparity_to_depth(d2, c, xy_offset=False, half=True)
    assert z2[5, 5] == 0 and z2[10, 10] == np.abs(np.float32(1) / (c0 + c1 * np.float32(600)))
    # the IR-grid offset (-4, -3): output columns 0, 1 and rows 0, 1 have sources outside the image
    z3 = osn.disparity_to_depth(d, c, xy_offset=True, half=True)
    assert np.all(z3[:, :2] == 0) and np.all(z3[:2, :] == 0) and z3[2, 2] == expect
    # normalised output, and out-of-window depth
    zn = osn.disparity_to_depth(d, c, xy_offset=False, half=True, mode=1)
    assert zn[3, 3] == (expect - mn) / (mx - mn)
    far = np.full((48, 64), 1080, np.uint16)  # 1 / (c0 + c1 d) beyond 7000 mm
    assert np.abs(np.float32(1) / (c0 + c1 * np.float32(1080))) > mx
    assert np.all(osn.disparity_to_depth(far, c, xy_offset=False)[:20, :20] == 0)
    assert np.all(osn.disparity_to_depth(far, c, xy_offset=False, mode=1)[:20, :20] == (np.float32(0) - mn) / (mx - mn))
    # rows at and below 4*floor(out_h/4) are never written
    z5 = osn.disparity_to_depth(np.full((46, 64), 800, np.uint16), c, xy_offset=False)  # out_h 23 -> rows 20.. zero
    assert np.all(z5[20:] == 0) and np.all(z5[:20] == expect)
    # full resolution
    zf = osn.disparity_to_depth(d, c, xy_offset=False, half=False)
    assert zf.shape == (48, 64) and np.all(zf[:47, :63] == expect)
    assert np.all(zf[:, 63] == 0) and np.all(zf[47] == 0)  # the last row and column of sources are excluded


def test_oracle_disparity_vectorised_equals_loops():
    rng = np.random.default_rng(5)
    c, _ = sensor.load_calib_xml(CALIB_XML, 320, 240)
    d = rng.integers(300, 1100, (40, 56)).astype(np.uint16)
    d[rng.random(d.shape) < 0.1] = 2047
    for half in (True, False):
        for off in (True, False):
            for mode in (0, 1):
                assert np.array_equal(osn.disparity_to_depth(d, c, off, half, mode),
                                      osn.disparity_to_depth_np(d, c, off, half, mode))
