"""HIP tracker vs the CPU oracle, kernel by kernel, through the C ABI.

Integer and index work (histograms, selection, compaction) and every per-pixel
or per-point FP32 value are compared bit for bit with the oracle's FP32
transcription (both sides compiled with -ffp-contract=off). The 27-term normal
equations are FP32 per workgroup and FP64 across workgroups on the device; they
are compared with the oracle's FP64 sums of the same per-point products within
the FP32 accumulation bound 2^-24 * 64 * sum|term| (row length 64 terms per
thread-level tree at most).
"""
import numpy as np
import pytest

from recon_amd import BIOBJECTIVE, GEOMETRIC, PHOTOMETRIC, synth
from recon_amd.tracker import PhotoTracker
from tests._oracle_steps import Steps

pytestmark = pytest.mark.gpu

CASES = [(320, 240, 3, 8192, 1), (640, 480, 4, 8192, 2), (160, 120, 3, 2048, 3)]
# the same stages under a two-camera rig calibration (distortion, TLR/TRL, KL != KR)
RIG_CASES = [(320, 240, 3, 8192, 5), (640, 480, 4, 8192, 6)]
CALIBS = {"pinhole": synth.calib_block_np, "rig": synth.calib_block_rig}


def _pair(W, H, seed, theta=2.0, t=30.0):
    return synth.make_pair(seed, W, H, theta, t)


def _setup(W, H, L, N, seed, mode=PHOTOMETRIC, calib="pinhole", **kw):
    g0, d0, g1, d1, Tgt = _pair(W, H, seed)
    cal = CALIBS[calib](W, H)
    tr = PhotoTracker(W, H, cal, mode=mode, levels=L, n_select=N, **kw)
    tr.set_reference_points(g0, d0, mode=mode, lam=0.5, force=True)
    tr.set_current_image(g1, d1)
    st = Steps(W, H, L, cal, g0, d0, g1, d1, nselect=N, dtype=np.float32)
    return tr, st, Tgt


def _trel_guess(Tgt, frac=0.5):
    """A pose part-way between identity and the truth (exercises real warps)."""
    from scipy.spatial.transform import Rotation
    r = Rotation.from_matrix(Tgt[:3, :3]).as_rotvec() * frac
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(r).as_matrix()
    T[:3, 3] = Tgt[:3, 3] * frac
    return T.astype(np.float32)


@pytest.mark.parametrize("W,H,L,N,seed", CASES)
def test_pyramids_bitexact(W, H, L, N, seed):
    print(f"case pyramids {W}x{H} seed {seed}", flush=True)
    tr, st, _ = _setup(W, H, L, N, seed)
    for l in range(L):
        n = (W >> l) * (H >> l)
        assert np.array_equal(tr.debug_read("REF_GRAY", l, np.float32, n), st.ref_gray[l]), f"ref level {l}"
        assert np.array_equal(tr.debug_read("CUR_GRAY", l, np.float32, n), st.cur_gray[l]), f"cur level {l}"
    assert np.array_equal(tr.debug_read("REF_Z", 0, np.float32, W * H), st.ref_z)
    assert np.array_equal(tr.debug_read("CUR_Z", 0, np.float32, W * H), st.cur_z)


def test_rgbd_texture_input_bitexact():
    """The reference's own float4 RGBD (bottom-up) input, rgb2gray.cu:64-76."""
    print("case rgbd 320x240", flush=True)
    W, H = 320, 240
    rng = np.random.default_rng(5)
    rgbd = rng.random((H, W, 4), dtype=np.float32)
    cal = synth.calib_block_np(W, H)
    tr = PhotoTracker(W, H, cal, mode=PHOTOMETRIC)
    tr.set_current_rgbd(rgbd)
    g = np.zeros(W * H, np.float32); z = np.zeros(W * H, np.float32)
    from oracle import call
    call("or_convert_rgbd_f", np.ascontiguousarray(rgbd), W, H, cal, g, z)
    assert np.array_equal(tr.debug_read("CUR_GRAY", 0, np.float32, W * H), g)
    assert np.array_equal(tr.debug_read("CUR_Z", 0, np.float32, W * H), z)


@pytest.mark.parametrize("W,H,L,N,seed,calib", [c + ("pinhole",) for c in CASES] + [c + ("rig",) for c in RIG_CASES])
def test_selection_bitexact(W, H, L, N, seed, calib):
    print(f"case selection {W}x{H} seed {seed} {calib}", flush=True)
    tr, st, _ = _setup(W, H, L, N, seed, calib=calib)
    tr.debug_prepare()
    st.features()
    assert np.array_equal(tr.debug_read("GRAD_MAG", 0, np.float32, W * H), st.gmag)
    assert np.array_equal(tr.debug_read("GRAD_HIST", 0, np.uint32, 256), st.hist)
    assert np.array_equal(tr.debug_read("SELECTED", 0, np.int32, N), st.sel)


@pytest.mark.parametrize("W,H,L,N,seed,calib", [c + ("pinhole",) for c in CASES] + [c + ("rig",) for c in RIG_CASES])
def test_points_and_jacobians_bitexact(W, H, L, N, seed, calib):
    print(f"case jacobians {W}x{H} seed {seed} {calib}", flush=True)
    tr, st, _ = _setup(W, H, L, N, seed, calib=calib)
    tr.debug_prepare()
    st.features()
    pts = tr.debug_read("POINTS", 0, np.float32, 4 * N).reshape(N, 4)
    rec = st.rec.reshape(-1, 24)[st.sel]
    assert np.array_equal(pts[:, :3], rec[:, :3])
    color_slot = [14, 17, 20, 23]
    for l in range(L):
        jac = tr.debug_read("JACOBIAN", l, np.float32, 7 * N)
        J = st.J[l].reshape(6, N).T
        got = np.concatenate([jac[:4 * N].reshape(N, 4), jac[4 * N:6 * N].reshape(N, 2)], axis=1)
        assert np.array_equal(got, J), f"jacobian level {l}"
        assert np.array_equal(jac[6 * N:], rec[:, color_slot[l]]), f"reference gray level {l}"


@pytest.mark.parametrize("W,H,L,N,seed,calib", [c + ("pinhole",) for c in CASES] + [c + ("rig",) for c in RIG_CASES])
@pytest.mark.parametrize("bins,robust", [(64, 0), (256, 0), (64, 1)])
def test_photometric_linearisation(W, H, L, N, seed, calib, bins, robust):
    print(f"case photo-lin {W}x{H} seed {seed} {calib} bins {bins} robust {robust}", flush=True)
    tr, st, Tgt = _setup(W, H, L, N, seed, calib=calib, mad_bins=bins, robust=robust)
    tr.debug_prepare()
    st.features()
    T = _trel_guess(Tgt)
    tr.debug_set_trel(T)
    for layer in range(L):
        (Ap, bp), _ = tr.debug_linearize(layer)
        o = st.photo_system(layer, T, bins=bins, robust=robust)
        rz = tr.debug_read("RESIDUAL", 0, np.float32, 2 * N).reshape(N, 2)
        assert np.array_equal(rz[:, 0], o["res"]), f"residual level {layer}"
        assert np.array_equal(rz[:, 1], o["zw"]), f"zWeight level {layer}"
        assert np.array_equal(tr.debug_read("MAD_HIST", 0, np.uint32, bins), o["hist"])
        tol = 64 * 2.0 ** -24
        assert np.all(np.abs(Ap - o["A"]) <= tol * o["Aabs"] + 1e-30), f"JtJ level {layer}"
        assert np.all(np.abs(bp - o["b"]) <= tol * o["babs"] + 1e-30), f"Jtr level {layer}"


@pytest.mark.parametrize("W,H,L,N,seed,calib", [c + ("pinhole",) for c in CASES[:2]] + [RIG_CASES[0] + ("rig",)])
def test_icp_clouds_bitexact(W, H, L, N, seed, calib):
    print(f"case clouds {W}x{H} seed {seed} {calib}", flush=True)
    tr, st, _ = _setup(W, H, L, N, seed, mode=GEOMETRIC, calib=calib)
    tr.debug_prepare()
    st.clouds()
    for l in range(L):
        n = (W >> l) * (H >> l)
        rc = tr.debug_read("REF_CLOUD", l, np.float32, 8 * n)
        pos, nrm = rc[:4 * n].reshape(n, 4), rc[4 * n:].reshape(n, 4)
        o = st.rpts[l].reshape(n, 6)
        m = st.rmask[l].astype(bool)
        assert np.array_equal(pos[:, 3] != 0, m), f"ref mask level {l}"
        assert np.array_equal(pos[m, :3], o[m, :3]), f"ref points level {l}"
        assert np.array_equal(nrm[m, :3], o[m, 3:]), f"ref normals level {l}"
        cc = tr.debug_read("CUR_CLOUD", l, np.float32, 4 * n).reshape(n, 4)
        oc = st.cpts[l].reshape(n, 6)
        mc = st.cmask[l].astype(bool)
        assert np.array_equal(cc[:, 3] != 0, mc), f"cur mask level {l}"
        assert np.array_equal(cc[mc, :3], oc[mc, :3]), f"cur points level {l}"


@pytest.mark.parametrize("W,H,L,N,seed,calib", [c + ("pinhole",) for c in CASES[:2]] + [RIG_CASES[0] + ("rig",)])
def test_icp_selection_weights_and_system(W, H, L, N, seed, calib):
    print(f"case icp {W}x{H} seed {seed} {calib}", flush=True)
    tr, st, Tgt = _setup(W, H, L, N, seed, mode=GEOMETRIC, calib=calib)
    tr.debug_prepare()
    st.clouds()
    T = _trel_guess(Tgt)
    tr.debug_set_trel(T)
    for layer in range(L):
        _, (Ag, bg) = tr.debug_linearize(layer, 0)
        o = st.icp_select(layer, T)
        n = int(tr.debug_read("ICP_COUNT", 0, np.int32, 4)[layer])
        assert n == o["n"], f"count level {layer}"
        nl = (W >> layer) * (H >> layer)
        assert np.array_equal(tr.debug_read("ICP_SELECTED", layer, np.int32, nl)[:n], o["sel"])
        assert np.array_equal(tr.debug_read("ICP_R2", layer, np.float32, nl)[:n], o["r2"])
        var = tr.debug_read("ICP_VAR", layer, np.float32, 3)
        assert var[2] / np.float32(n) == o["sigma2"], f"student-t variance level {layer}"
        s = st.icp_system(layer, T, o["sel"], o["w"])
        tol = 64 * 2.0 ** -24
        assert np.all(np.abs(Ag - s["A"]) <= tol * s["Aabs"] + 1e-30), f"icp JtJ level {layer}"
        assert np.all(np.abs(bg - s["b"]) <= tol * s["babs"] + 1e-30), f"icp Jtr level {layer}"


def _rot_deg(T):
    c = np.clip((np.trace(np.asarray(T, np.float64)[:3, :3]) - 1) / 2, -1, 1)
    return np.degrees(np.arccos(c))


@pytest.mark.parametrize("mode", [PHOTOMETRIC, GEOMETRIC, BIOBJECTIVE])
@pytest.mark.parametrize("solver", [0, 1])
def test_track_matches_oracle(mode, solver):
    """Full coarse-to-fine tracking of a 320x240 pair: the device pose against
    the FP32 oracle (same arithmetic, different summation order) and the FP64
    oracle, and both against the ground truth."""
    import oracle
    print(f"case track mode {mode} solver {solver}", flush=True)
    W, H, L, N = 320, 240, 3, 8192
    g0, d0, g1, d1, Tgt = _pair(W, H, 11)
    cal = synth.calib_block_np(W, H)
    tr = PhotoTracker(W, H, cal, mode=mode, levels=L, n_select=N, solver=solver, iters=(10, 10, 10, 0))
    tr.set_reference_points(g0, d0, mode=mode, lam=0.5, force=True)
    tr.set_current_image(g1, d1)
    assert tr.track()
    Tdev = tr.trel().astype(np.float64)
    res = {}
    for dt in (np.float32, np.float64):
        o = oracle.Tracker(W, H, cal, mode=mode, levels=L, iters=(10, 10, 10, 0), nselect=N, solver=solver,
                           lam=0.5, dtype=dt)
        o.set_reference(g0, d0); o.set_current(g1, d1); o.track()
        res[dt] = o.poses()[0]
    for dt, To in res.items():
        rot, tr_mm = synth.pose_error(Tdev, To)
        print(f"  device vs oracle {dt.__name__}: {rot:.2e} deg, {tr_mm:.2e} mm")
        if dt == np.float32:
            # the reference's arithmetic restated: agreement within 1e-5 rad / 1e-2 mm
            assert np.radians(rot) < 1e-5 and tr_mm < 1e-2
        else:
            # FP64 oracle: the FP32 path's own rounding, stated per result
            assert rot < 1e-2 and tr_mm < 0.1
    rot, tmm = synth.pose_error(Tdev, Tgt)
    print(f"  device vs ground truth: {rot:.3f} deg, {tmm:.3f} mm (start {_rot_deg(Tgt):.3f} deg)")
    assert rot < 0.1 and tmm < 5.0
    # host pose bookkeeping (applyIncrement): T is the inverse of Trel
    assert np.allclose(tr.get_relative_pose().astype(np.float64) @ Tdev, np.eye(4), atol=1e-4)
    assert np.allclose(tr.get_global_pose(), np.linalg.inv(Tdev), atol=1e-3)


@pytest.mark.parametrize("W,H,L,N,seed", CASES[:2])
def test_inspection_images_bitexact(W, H, L, N, seed):
    """updateErrorImage / getDebugImage1C / getRefImage1C against
    oracle/debug_images.py on the oracle's own intermediate results, at a pose
    part-way to the truth: photometric error at the selected pixels, the
    keyframe's normalised depth, the geometric error of every level, and the
    geometric reference images the reference never writes."""
    from oracle import debug_images as di
    print(f"case inspection images {W}x{H} seed {seed}", flush=True)
    tr, st, Tgt = _setup(W, H, L, N, seed, mode=PHOTOMETRIC)
    tr.debug_prepare()
    st.features()
    T = _trel_guess(Tgt)
    tr.debug_set_trel(T)
    tr.update_error_image()
    res, _ = st.photo_residual(0, T)
    exp = di.photo_error_image(st.sel, res, W * H)
    assert np.array_equal(tr.get_debug_image().ravel(), exp)
    assert np.array_equal(tr.get_ref_image().ravel(), di.ref_depth_image(st.ref_z, st.calib))
    tr, st, Tgt = _setup(W, H, L, N, seed, mode=GEOMETRIC)
    tr.debug_prepare()
    st.clouds()
    tr.debug_set_trel(T)
    tr.update_error_image()
    for l in range(L):
        w, h = W >> l, H >> l
        e = di.icp_error_image(st.rpts[l], st.rmask[l], st.cpts[l], st.cmask[l], w, h, T, st.calib, l)
        got = tr.get_debug_image(l).ravel()
        assert np.array_equal(got, e), f"geometric error image level {l}"
        assert (got > 0).mean() > 0.2
        assert np.all(tr.get_ref_image(l) == 0)
