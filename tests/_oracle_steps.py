"""Oracle intermediate results for one keyframe/current pair, step by step
(the per-kernel restatements of oracle/oracle.hpp), in FP32 or FP64."""
import ctypes

import numpy as np

import oracle
from oracle import call


def _t(dtype):
    return "f" if dtype == np.float32 else "d"


class Steps:
    def __init__(self, W, H, L, calib, g0, d0, g1, d1, nselect=8192, dtype=np.float32):
        self.W, self.H, self.L, self.N = W, H, L, nselect
        self.dt = dtype
        s = _t(dtype)
        self.s = s
        self.calib = np.ascontiguousarray(calib, np.float32)
        n = W * H
        self.ref_gray = [np.zeros(n, dtype)]
        self.ref_z = np.zeros(n, dtype)
        call(f"or_convert_u8u16_{s}", np.ascontiguousarray(g0, np.uint8), np.ascontiguousarray(d0, np.uint16), n,
             self.ref_gray[0], self.ref_z)
        self.cur_gray = [np.zeros(n, dtype)]
        self.cur_z = np.zeros(n, dtype)
        call(f"or_convert_u8u16_{s}", np.ascontiguousarray(g1, np.uint8), np.ascontiguousarray(d1, np.uint16), n,
             self.cur_gray[0], self.cur_z)
        for pyr in (self.ref_gray, self.cur_gray):
            for l in range(1, L):
                lo = np.zeros((W >> l) * (H >> l), dtype)
                call(f"or_downsample2_{s}", pyr[l - 1], W >> (l - 1), lo, W >> l, H >> l)
                pyr.append(lo)

    def features(self):
        W, H, s = self.W, self.H, self.s
        self.rec = np.zeros(W * H * 24, self.dt)
        call(f"or_vertex_records_{s}", self.ref_z, self.ref_gray[0], W, H, self.calib, self.rec)
        self.gmag = np.ascontiguousarray(self.rec.reshape(-1, 24)[:, 13])
        self.hist = np.zeros(256, np.uint32)
        call(f"or_grad_histogram_{s}", self.gmag, W * H, 256, self.hist)
        self.sel = np.zeros(self.N, np.int32)
        thr = np.zeros(1, np.int32); exc = np.zeros(1, np.int32)
        call(f"or_select_points_{s}", self.gmag, W * H, self.hist, 256, self.N, self.sel, thr, exc)
        self.thr, self.excess = int(thr[0]), int(exc[0])
        g = self.ref_gray + [None] * (4 - self.L)
        g = [x if x is not None else self.ref_gray[0] for x in g]
        call(f"or_point_attributes_{s}", self.sel, self.N, self.rec, g[0], g[1], g[2], g[3], W, H, self.L, self.calib)
        self.J = [np.zeros(6 * self.N, self.dt) for _ in range(4)]
        call(f"or_jacobians_{s}", self.sel, self.N, self.rec, self.calib, ctypes.c_float(1e-3), self.L, *self.J)
        return self

    def photo_residual(self, layer, Trel):
        T = np.ascontiguousarray(Trel, self.dt).ravel()
        res = np.zeros(self.N, self.dt); zw = np.zeros(self.N, self.dt)
        call(f"or_photo_residual_{self.s}", self.sel, self.N, self.rec, T, self.calib, layer, self.cur_gray[layer],
             self.W >> layer, self.H >> layer, self.cur_z, self.W, self.H, res, zw)
        return res, zw

    def photo_system(self, layer, Trel, bins=64, robust=0):
        res, zw = self.photo_residual(layer, Trel)
        hist = np.zeros(bins, np.uint32)
        m = call(f"or_mad_median_bin_{self.s}", res, self.N, bins, hist)
        w = np.zeros(self.N, self.dt); wr = np.zeros(self.N, self.dt)
        call(f"or_robust_weights_{self.s}", res, zw, self.N, m, bins, robust, w, wr)
        A = np.zeros(36); b = np.zeros(6)
        call(f"or_normal_equations_{self.s}", self.J[layer], w, wr, self.N, A, b)
        # per-term magnitude bound for FP32 accumulation error
        Jw = self.J[layer].reshape(6, -1).astype(np.float64) * w.astype(np.float64)
        Aabs = np.abs(Jw) @ np.abs(Jw).T
        babs = np.abs(Jw) @ np.abs(wr.astype(np.float64))
        return dict(res=res, zw=zw, hist=hist, m=m, w=w, A=A.reshape(6, 6), b=b, Aabs=Aabs, babs=babs)

    # ---- geometric side
    def clouds(self):
        W, H, s = self.W, self.H, self.s
        self.rpts = [np.zeros(W * H * 6, self.dt)]
        self.rmask = [np.zeros(W * H, np.int32)]
        call(f"or_ref_cloud_{s}", self.ref_z, W, H, self.calib, self.rpts[0], self.rmask[0])
        self.cpts = [np.zeros(W * H * 6, self.dt)]
        self.cmask = [np.zeros(W * H, np.int32)]
        call(f"or_cur_cloud_{s}", self.cur_z, W, H, self.calib, self.cpts[0], self.cmask[0])
        for l in range(1, self.L):
            n = (W >> l) * (H >> l)
            for pts, msk in ((self.rpts, self.rmask), (self.cpts, self.cmask)):
                p = np.zeros(n * 6, self.dt); m = np.zeros(n, np.int32)
                call(f"or_downsample_cloud_{s}", pts[l - 1], msk[l - 1], W >> l, H >> l, p, m)
                pts.append(p); msk.append(m)
        return self

    def icp_select(self, layer, Trel):
        w, h = self.W >> layer, self.H >> layer
        T = np.ascontiguousarray(Trel, self.dt).ravel()
        full = np.zeros(w * h, self.dt); rm = np.zeros(w * h, np.int32)
        call(f"or_icp_residual_mask_{self.s}", self.rpts[layer], self.rmask[layer], w, h, T, self.calib,
             self.cpts[layer], self.cmask[layer], layer, full, rm)
        sel = np.zeros(w * h, np.int32)
        n = call("or_compact", rm, w * h, sel)
        sel = sel[:n]
        r2 = (full[sel] * full[sel]).astype(self.dt)
        wts = np.zeros(n, self.dt); sig = np.zeros(1, self.dt)
        call(f"or_student_t_{self.s}", r2, n, wts, sig)
        return dict(full=full, mask=rm, sel=sel, n=n, r2=r2, w=wts, sigma2=sig[0])

    def icp_system(self, layer, Trel, sel, wts):
        w, h = self.W >> layer, self.H >> layer
        n = len(sel)
        T = np.ascontiguousarray(Trel, self.dt).ravel()
        res = np.zeros(n, self.dt); J = np.zeros(6 * n, self.dt)
        call(f"or_icp_residual_jacobian_{self.s}", self.rpts[layer], np.ascontiguousarray(sel, np.int32), n, w, h,
             layer, T, self.calib, self.cpts[layer], self.cmask[layer], np.ascontiguousarray(wts, self.dt),
             ctypes.c_float(1e-3), res, J)
        ones = np.ones(n, self.dt)
        A = np.zeros(36); b = np.zeros(6)
        call(f"or_normal_equations_{self.s}", J, ones, res, n, A, b)
        Jd = np.abs(J.reshape(6, -1).astype(np.float64))
        return dict(res=res, J=J, A=A.reshape(6, 6), b=b, Aabs=Jd @ Jd.T, babs=Jd @ np.abs(res.astype(np.float64)))
