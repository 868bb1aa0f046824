// CPU oracle: the tracker driver (TEST INFRASTRUCTURE ONLY, see oracle.hpp).
//
// Restates the per-frame flow of PhotoTracker / ShapeTracker
// (tracker/tracker/phototracker.cpp:360-386 prepareReference, :505-512
// updateLinearSystem, :646-664 estimateIncrement, :666-689 gpuEstimateIncrement,
// :514-531 applyIncrement; tracker/tracker/shapetracker.cpp:195-211
// setReferencePoints, :254-288 pyramids/prepare, :357-382 updateLinearSystem)
// over the per-kernel restatements of oracle.hpp, plus the keyframe fusion rule
// of the build (a per-pixel running weighted mean, the TSDF update rule of
// tracker/kernels/VoxelGrid_cuda.cu:1011-1013 applied to keyframe depth and gray).
#pragma once
#include "oracle.hpp"

namespace oracle {

enum { MODE_GEOMETRIC = 0, MODE_PHOTOMETRIC = 1, MODE_BIOBJECTIVE = 2 };
enum { SOLVER_CHOLESKY = 0, SOLVER_CG = 1 };

struct Params {
    int mode = MODE_PHOTOMETRIC;
    float lambda = 0.9f;
    int levels = 0;             // 0: the reference rule (4 at width 640, else 3)
    int iters[4] = {1, 3, 10, 5};  // coarse -> fine
    int nselect = 8192;
    int mad_bins = 64;
    int robust = ROBUST_TUKEY;
    int solver = SOLVER_CHOLESKY;
    float optScaleIn = 1e-3f, optScaleOut = 1e3f;
};

inline int default_levels(int W) { return W == 640 ? 4 : 3; }  // phototracker.cpp:777-784

template <class R, class P>
void fuse_frame(R* kz, R* kg, R* kw, const R* fz, const R* fg, int W, int H, const P* Trel, const float* calib) {
    const float* KL = calib + KL_OFFSET;
    R mn = (R)calib[MIND_OFFSET], mx = (R)calib[MAXD_OFFSET];
    P iT[16];
    invert_rt4<P>(Trel, iT);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            size_t o = (size_t)x + (size_t)y * W;
            R p4[4];
            backproject<R>(kz[o], x, y, calib, p4);
            if (p4[3] == 0) continue;
            V3<R> q = mul4<R>(Trel, V3<R>{p4[0], p4[1], p4[2]});
            R u = (R)KL[0] * (q.x / q.z) + (R)KL[2];
            R v = (R)KL[4] * (q.y / q.z) + (R)KL[5];
            if (!(u >= (R)0 && v >= (R)0 && u < (R)(W - 1) && v < (R)(H - 1))) continue;
            int xi = (int)(u + (R)0.5f), yi = (int)(v + (R)0.5f);
            size_t fo = (size_t)xi + (size_t)yi * W;
            R zf = fz[fo];
            if (!(zf < -mn && zf > -mx)) continue;
            R dz = zf - q.z;
            if (!(std::fabs(dz) < (R)30.0f)) continue;
            R f4[4];
            backproject<R>(zf, xi, yi, calib, f4);
            V3<R> pk = mul4<R>(iT, V3<R>{f4[0], f4[1], f4[2]});
            int xd = (int)u, yd = (int)v;
            R gobs = bilerp<R>(fg, W, xd, yd, u - (R)xd, v - (R)yd);
            R Wold = kw[o];
            R Wn = Wold + (R)1;
            kz[o] = (Wold * kz[o] + pk.z) / Wn;
            kg[o] = (Wold * kg[o] + gobs) / Wn;
            kw[o] = std::min(Wn, (R)16);
        }
}


template <class R, class P>
struct Tracker {
    using Vec = std::vector<R>;
    int W = 0, H = 0, L = 3;
    Params prm;
    float calib[CALIB_SIZE];
    // keyframe
    Vec kf_gray0, kf_z;           // raw keyframe images (after conversion / fusion)
    Vec kf_fw;                     // fusion weights
    Vec ref_gray[4];
    Vec rec;                       // W*H*24 vertex records
    std::vector<int> sel;          // selected indices
    Vec Jl[4];                     // 6*N per level
    Vec ref_cloud_[4]; std::vector<int> ref_mask[4];
    // current
    Vec cur_gray[4], cur_z;
    Vec cur_cloud_[4]; std::vector<int> cur_mask[4];
    // pose state (phototracker.h members)
    P Trel[16];          // TrelDev: maps reference points into the current view
    float T[16];         // host relative pose (inverse of Trel)
    double Tinc[16], global[16];
    bool ref_prepared = false, icp_prepared = false;
    bool pending_pose_reset = false;  // a new keyframe clears the relative pose when prepared
    // per-layer ICP state
    std::vector<int> icp_sel[4]; Vec icp_w[4]; int nsel_icp[4] = {0, 0, 0, 0};
    // diagnostics of the last linearisation
    double lastA[36], lastb[6];
    int last_median = 0;

    void init(int w, int h, const float* cal, const Params& p) {
        W = w; H = h; prm = p;
        L = p.levels > 0 ? p.levels : default_levels(w);
        std::memcpy(calib, cal, sizeof(calib));
        for (int i = 0; i < 16; i++) { Trel[i] = (i % 5 == 0); T[i] = (i % 5 == 0); Tinc[i] = (i % 5 == 0); global[i] = (i % 5 == 0); }
    }
    int wl(int l) const { return W >> l; }
    int hl(int l) const { return H >> l; }

    void build_pyramid(Vec* pyr) {
        for (int l = 1; l < L; l++) {
            pyr[l].assign((size_t)wl(l) * hl(l), 0);
            downsample2<R>(pyr[l - 1].data(), wl(l - 1), pyr[l].data(), wl(l), hl(l));
        }
    }
    void set_reference_u8(const uint8_t* g, const uint16_t* d) {
        kf_gray0.assign((size_t)W * H, 0); kf_z.assign((size_t)W * H, 0);
        convert_u8u16<R>(g, d, W * H, kf_gray0.data(), kf_z.data());
        kf_fw.assign((size_t)W * H, (R)1);
        ref_prepared = false; icp_prepared = false;
        pending_pose_reset = true;
        set_icp_reference();
    }
    // setReferencePoints with ray-cast maps and the filtered reference on
    // (phototracker.cpp:310-337, shapetracker.cpp:199-200): the photometric
    // keyframe from the frame, the ICP cloud from the maps
    void set_reference_points_u8(const uint8_t* g, const uint16_t* d, const float* xyz, const float* nrm) {
        set_reference_u8(g, d);
        ref_cloud_[0].assign((size_t)W * H * 6, 0); ref_mask[0].assign((size_t)W * H, 0);
        oriented_cloud<R>(xyz, nrm, W, H, ref_cloud_[0].data(), ref_mask[0].data());
    }
    // ShapeTracker::setReferencePoints, raw-measurement branch (shapetracker.cpp:195-211)
    void set_icp_reference() {
        ref_cloud_[0].assign((size_t)W * H * 6, 0); ref_mask[0].assign((size_t)W * H, 0);
        ref_cloud<R>(kf_z.data(), W, H, calib, ref_cloud_[0].data(), ref_mask[0].data());
    }
    void set_current_u8(const uint8_t* g, const uint16_t* d) {
        cur_gray[0].assign((size_t)W * H, 0); cur_z.assign((size_t)W * H, 0);
        convert_u8u16<R>(g, d, W * H, cur_gray[0].data(), cur_z.data());
    }
    // PhotoTracker::prepareReference (phototracker.cpp:360-386) + selectPixels (:343-350)
    void reset_relative_pose() {
        for (int i = 0; i < 16; i++) Trel[i] = (i % 5 == 0);
        for (int i = 0; i < 16; i++) T[i] = (i % 5 == 0);
    }
    void prepare_photo_reference() {
        if (pending_pose_reset) reset_relative_pose();
        ref_gray[0] = kf_gray0;
        build_pyramid(ref_gray);
        rec.assign((size_t)W * H * VB_STRIDE, 0);
        vertex_records<R>(kf_z.data(), ref_gray[0].data(), W, H, calib, rec.data());
        Vec g((size_t)W * H);
        for (size_t i = 0; i < g.size(); i++) g[i] = rec[i * VB_STRIDE + VB_GMAG];
        std::vector<uint32_t> hist(256);
        grad_histogram<R>(g.data(), W * H, 256, hist.data());
        sel.assign(prm.nselect, 0);
        select_points<R>(g.data(), W * H, hist.data(), 256, prm.nselect, sel.data());
        const R* pyr[4] = {ref_gray[0].data(), L > 1 ? ref_gray[1].data() : nullptr,
                           L > 2 ? ref_gray[2].data() : nullptr, L > 3 ? ref_gray[3].data() : nullptr};
        point_attributes<R>(sel.data(), prm.nselect, rec.data(), pyr, W, H, L, calib);
        R* Jp[4];
        for (int l = 0; l < 4; l++) { Jl[l].assign((size_t)6 * prm.nselect, 0); Jp[l] = Jl[l].data(); }
        jacobians<R>(sel.data(), prm.nselect, rec.data(), calib, prm.optScaleIn, L, Jp);
        ref_prepared = true;
    }
    // ShapeTracker::prepareReference (shapetracker.cpp:273-288)
    void prepare_icp_reference() {
        if (pending_pose_reset) reset_relative_pose();
        for (int l = 1; l < L; l++) {
            ref_cloud_[l].assign((size_t)wl(l) * hl(l) * 6, 0); ref_mask[l].assign((size_t)wl(l) * hl(l), 0);
            downsample_cloud<R>(ref_cloud_[l - 1].data(), ref_mask[l - 1].data(), wl(l), hl(l), ref_cloud_[l].data(), ref_mask[l].data());
        }
        icp_prepared = true;
    }
    void prepare_current_photo() { build_pyramid(cur_gray); }
    void prepare_current_icp() {
        cur_cloud_[0].assign((size_t)W * H * 6, 0); cur_mask[0].assign((size_t)W * H, 0);
        cur_cloud<R>(cur_z.data(), W, H, calib, cur_cloud_[0].data(), cur_mask[0].data());
        for (int l = 1; l < L; l++) {
            cur_cloud_[l].assign((size_t)wl(l) * hl(l) * 6, 0); cur_mask[l].assign((size_t)wl(l) * hl(l), 0);
            downsample_cloud<R>(cur_cloud_[l - 1].data(), cur_mask[l - 1].data(), wl(l), hl(l), cur_cloud_[l].data(), cur_mask[l].data());
        }
    }
    // PhotoTracker::updateLinearSystem (phototracker.cpp:505-512)
    void photo_system(int layer, double A[36], double b[6]) {
        int n = prm.nselect;
        Vec res(n), zw(n), w(n), wr(n);
        photo_residual<R, P>(sel.data(), n, rec.data(), Trel, calib, layer, cur_gray[layer].data(), wl(layer), hl(layer),
                             cur_z.data(), W, H, res.data(), zw.data());
        int m = mad_median_bin<R>(res.data(), n, prm.mad_bins);
        last_median = m;
        robust_weights<R>(res.data(), zw.data(), n, m, prm.mad_bins, prm.robust, w.data(), wr.data());
        const R* rows[6];
        for (int k = 0; k < 6; k++) rows[k] = Jl[layer].data() + (size_t)k * n;
        normal_equations<R>(rows, w.data(), wr.data(), n, A, b);
    }
    // ShapeTracker::updateLinearSystem (shapetracker.cpp:357-382)
    bool icp_system(int layer, int iteration, double A[36], double b[6]) {
        int w = wl(layer), h = hl(layer);
        if (iteration == 0) {
            Vec full((size_t)w * h, 0);
            std::vector<int> rm((size_t)w * h, 0);
            icp_residual_mask<R, P>(ref_cloud_[layer].data(), ref_mask[layer].data(), w, h, Trel, calib,
                                    cur_cloud_[layer].data(), cur_mask[layer].data(), layer, full.data(), rm.data());
            icp_sel[layer].assign((size_t)w * h, 0);
            nsel_icp[layer] = compact(rm.data(), w * h, icp_sel[layer].data());
            int n = nsel_icp[layer];
            if (n <= 0) return false;
            Vec r2(n);
            for (int i = 0; i < n; i++) { R r = full[icp_sel[layer][i]]; r2[i] = r * r; }
            icp_w[layer].assign(n, 0);
            student_t_weights<R>(r2.data(), n, icp_w[layer].data());
        }
        int n = nsel_icp[layer];
        if (n <= 0) return false;
        Vec res(n), J(6 * (size_t)n);
        R* Jp[6];
        for (int k = 0; k < 6; k++) Jp[k] = J.data() + (size_t)k * n;
        icp_residual_jacobian<R, P>(ref_cloud_[layer].data(), icp_sel[layer].data(), n, w, h, layer, Trel, calib,
                                    cur_cloud_[layer].data(), cur_mask[layer].data(), icp_w[layer].data(),
                                    prm.optScaleIn, res.data(), Jp);
        Vec ones(n, (R)1);
        const R* rows[6];
        for (int k = 0; k < 6; k++) rows[k] = Jp[k];
        normal_equations<R>(rows, ones.data(), res.data(), n, A, b);
        return true;
    }
    // PhotoTracker::estimateIncrement (phototracker.cpp:646-664) + solveMotionCuda (warp.cu:1512-1601)
    void solve_and_update(double A[36], double b[6]) {
        double x[6];
        if (prm.solver == SOLVER_CG) solve_cg(A, b, x);
        else solve_cholesky(A, b, x);
        std::memcpy(lastA, A, sizeof(lastA));
        std::memcpy(lastb, b, sizeof(lastb));
        expm_update<P>(x, Trel, (double)prm.optScaleOut);
    }
    // gpuPreProcess + gpuEstimateIncrement + applyIncrement (phototracker.cpp:596-689, 514-531)
    void track() {
        int mode = prm.mode;
        if (mode == MODE_GEOMETRIC || mode == MODE_BIOBJECTIVE) {
            if (!icp_prepared) prepare_icp_reference();
            prepare_current_icp();
        }
        if (mode == MODE_PHOTOMETRIC || mode == MODE_BIOBJECTIVE) {
            if (!ref_prepared) prepare_photo_reference();
            prepare_current_photo();
        }
        pending_pose_reset = false;
        for (int layer = L - 1; layer >= 0; layer--) {
            int n = prm.iters[L - 1 - layer];
            for (int j = 0; j < n; j++) {
                double Ap[36], bp[6], Ag[36], bg[6];
                bool ok = true;
                if (mode == MODE_PHOTOMETRIC || mode == MODE_BIOBJECTIVE) photo_system(layer, Ap, bp);
                if (mode == MODE_GEOMETRIC || mode == MODE_BIOBJECTIVE) ok = icp_system(layer, j, Ag, bg);
                if (!ok) continue;
                if (mode == MODE_PHOTOMETRIC) solve_and_update(Ap, bp);
                else if (mode == MODE_GEOMETRIC) solve_and_update(Ag, bg);
                else {
                    linear_fuse(Ap, bp, prm.nselect, Ag, bg, nsel_icp[layer], (double)prm.lambda);
                    solve_and_update(Ap, bp);
                }
            }
        }
        float Tf[16];
        for (int i = 0; i < 16; i++) Tf[i] = (float)Trel[i];
        apply_increment(Tf, T, Tinc, global);
    }
    // Keyframe fusion: every valid keyframe pixel is back-projected, moved into the
    // tracked frame by Trel, and matched to the nearest frame pixel; a depth within
    // 30 mm of the prediction updates the keyframe depth (the observed point moved
    // back into the keyframe) and gray by a running weighted mean, weight cap 16.
    // The fused keyframe takes effect when refresh_keyframe() re-prepares it; the
    // relative pose is kept.
    void fuse_current() {
        fuse_frame<R, P>(kf_z.data(), kf_gray0.data(), kf_fw.data(), cur_z.data(), cur_gray[0].data(), W, H, Trel, calib);
    }
    void refresh_keyframe() {
        ref_prepared = false; icp_prepared = false;
        set_icp_reference();
    }
};

}  // namespace oracle
