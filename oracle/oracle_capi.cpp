// C entry points of the CPU oracle (TEST INFRASTRUCTURE ONLY, see oracle.hpp),
// loaded by the tests through ctypes. Suffix _f: FP32 literal transcription,
// _d: FP64 oracle.
#include "cnn.hpp"
#include "tracker.hpp"
#include "tsdf.hpp"

using namespace oracle;

#define OR_BOTH(MACRO) MACRO(f, float) MACRO(d, double)

extern "C" {

#define DEF_CONVERT(S, R)                                                                             \
    void or_convert_u8u16_##S(const uint8_t* g, const uint16_t* d, int n, R* gray, R* z) {           \
        convert_u8u16<R>(g, d, n, gray, z);                                                           \
    }                                                                                                 \
    void or_convert_rgbd_##S(const float* rgbd, int W, int H, const float* calib, R* gray, R* z) {    \
        convert_rgbd<R>(rgbd, W, H, calib, gray, z);                                                  \
    }                                                                                                 \
    void or_downsample2_##S(const R* hi, int Wh, R* lo, int Wl, int Hl) { downsample2<R>(hi, Wh, lo, Wl, Hl); } \
    void or_vertex_records_##S(const R* z, const R* gray0, int W, int H, const float* calib, R* rec) {          \
        vertex_records<R>(z, gray0, W, H, calib, rec);                                                \
    }                                                                                                 \
    void or_grad_histogram_##S(const R* g, int n, int nbins, uint32_t* hist) {                        \
        grad_histogram<R>(g, n, nbins, hist);                                                         \
    }                                                                                                 \
    void or_select_points_##S(const R* g, int n, const uint32_t* hist, int nbins, int nselect, int* out, \
                              int* thr, int* excess) {                                                \
        select_points<R>(g, n, hist, nbins, nselect, out, thr, excess);                               \
    }                                                                                                 \
    void or_point_attributes_##S(const int* idx, int n, R* rec, const R* g0, const R* g1, const R* g2, \
                                 const R* g3, int W, int H, int L, const float* calib) {             \
        const R* g[4] = {g0, g1, g2, g3};                                                             \
        point_attributes<R>(idx, n, rec, g, W, H, L, calib);                                          \
    }                                                                                                 \
    void or_jacobians_##S(const int* idx, int n, const R* rec, const float* calib, float optScaleIn, int L, \
                          R* J0, R* J1, R* J2, R* J3) {                                               \
        R* J[4] = {J0, J1, J2, J3};                                                                   \
        jacobians<R>(idx, n, rec, calib, optScaleIn, L, J);                                           \
    }                                                                                                 \
    void or_photo_residual_##S(const int* idx, int n, const R* rec, const R* T, const float* calib, int layer, \
                               const R* img, int wl, int hl, const R* zcur, int W, int H, R* res, R* zw) { \
        photo_residual<R, R>(idx, n, rec, T, calib, layer, img, wl, hl, zcur, W, H, res, zw);        \
    }                                                                                                 \
    int or_mad_median_bin_##S(const R* res, int n, int bins, uint32_t* hist) {                        \
        return mad_median_bin<R>(res, n, bins, hist);                                                 \
    }                                                                                                 \
    void or_robust_weights_##S(const R* res, const R* zw, int n, int m, int bins, int robust, R* w, R* wr) { \
        robust_weights<R>(res, zw, n, m, bins, robust, w, wr);                                        \
    }                                                                                                 \
    void or_normal_equations_##S(const R* J, const R* w, const R* wr, int n, double* A, double* b) {  \
        const R* rows[6];                                                                             \
        for (int k = 0; k < 6; k++) rows[k] = J + (size_t)k * n;                                      \
        normal_equations<R>(rows, w, wr, n, A, b);                                                    \
    }                                                                                                 \
    void or_ref_cloud_##S(const R* z, int W, int H, const float* calib, R* pts, int* mask) {          \
        ref_cloud<R>(z, W, H, calib, pts, mask);                                                      \
    }                                                                                                 \
    void or_cur_cloud_##S(const R* z, int W, int H, const float* calib, R* pts, int* mask) {          \
        cur_cloud<R>(z, W, H, calib, pts, mask);                                                      \
    }                                                                                                 \
    void or_downsample_cloud_##S(const R* src, const int* msrc, int Wd, int Hd, R* dst, int* mdst) {  \
        downsample_cloud<R>(src, msrc, Wd, Hd, dst, mdst);                                            \
    }                                                                                                 \
    void or_icp_residual_mask_##S(const R* ref, const int* rmask, int w, int h, const R* T, const float* calib, \
                                  const R* cur, const int* cmask, int layer, R* res, int* rm) {       \
        icp_residual_mask<R, R>(ref, rmask, w, h, T, calib, cur, cmask, layer, res, rm);              \
    }                                                                                                 \
    void or_student_t_##S(const R* r2, int n, R* w, R* sigma2) { student_t_weights<R>(r2, n, w, sigma2); } \
    void or_icp_residual_jacobian_##S(const R* ref, const int* sel, int n, int w, int h, int layer, const R* T, \
                                      const float* calib, const R* cur, const int* cmask, const R* wts,        \
                                      float optScaleIn, R* res, R* J) {                               \
        R* Jp[6];                                                                                     \
        for (int k = 0; k < 6; k++) Jp[k] = J + (size_t)k * n;                                        \
        icp_residual_jacobian<R, R>(ref, sel, n, w, h, layer, T, calib, cur, cmask, wts, optScaleIn, res, Jp); \
    }                                                                                                 \
    void or_fuse_frame_##S(R* kz, R* kg, R* kw, const R* fz, const R* fg, int W, int H, const R* T,    \
                           const float* calib) {                                                      \
        fuse_frame<R, R>(kz, kg, kw, fz, fg, W, H, T, calib);                                         \
    }
OR_BOTH(DEF_CONVERT)

int or_compact(const int* mask, int n, int* out) { return compact(mask, n, out); }
void or_solve_cg(const double* A, const double* b, double* x) { solve_cg(A, b, x); }
int or_solve_cholesky(const double* A, const double* b, double* x) { return solve_cholesky(A, b, x) ? 1 : 0; }
void or_se3_exp(const double* x, double* E) { se3_exp(x, E); }
void or_normalize_mat4_safe(double* T) { normalize_mat4_safe(T); }
void or_expm_update_f(const double* x, float* T, double scale) { expm_update<float>(x, T, scale); }
void or_expm_update_d(const double* x, double* T, double scale) { expm_update<double>(x, T, scale); }
void or_apply_increment(const float* Trel, float* T, double* Tinc, double* global) {
    apply_increment(Trel, T, Tinc, global);
}
void or_rot2quat(const float* m, float* q, float* qn) {
    rot2quat(m, q);
    float len = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + 1e-7f;
    len = (float)std::sqrt((double)len);
    for (int i = 0; i < 4; i++) qn[i] = q[i] / len;
}
void or_invert_rt4_d(const double* M, double* Mi) { invert_rt4<double>(M, Mi); }
void or_oriented_cloud_f(const float* xyz, const float* nrm, int W, int H, float* pts, int* mask) {
    oriented_cloud<float>(xyz, nrm, W, H, pts, mask);
}
void or_invert_rt4_f(const float* M, float* Mi) { invert_rt4<float>(M, Mi); }
void or_mat4_mul(const double* A, const double* B, double* C) { mat4_mul(A, B, C); }
int or_pose_difference_exceeded(const float* dT, float tthr, float athr) {
    return pose_difference_exceeded(dT, tthr, athr) ? 1 : 0;
}
void or_linear_fuse(double* A1, double* b1, int N1, const double* A2, const double* b2, int N2, double lambda) {
    linear_fuse(A1, b1, N1, A2, b2, N2, lambda);
}

// ---- tracker objects. prm: {mode, levels, iters0..3, nselect, mad_bins, robust, solver}, lambda separate
static Params make_params(const int* p, float lambda) {
    Params q;
    q.mode = p[0]; q.levels = p[1];
    for (int i = 0; i < 4; i++) q.iters[i] = p[2 + i];
    q.nselect = p[6]; q.mad_bins = p[7]; q.robust = p[8]; q.solver = p[9];
    q.lambda = lambda;
    return q;
}

#define DEF_TRACKER(S, R)                                                                            \
    typedef Tracker<R, R> Tr_##S;                                                                    \
    void* or_tracker_create_##S(int W, int H, const float* calib, const int* prm, float lambda) {    \
        Tr_##S* t = new Tr_##S();                                                                    \
        t->init(W, H, calib, make_params(prm, lambda));                                              \
        return t;                                                                                    \
    }                                                                                                \
    void or_tracker_destroy_##S(void* h) { delete (Tr_##S*)h; }                                      \
    void or_tracker_set_reference_##S(void* h, const uint8_t* g, const uint16_t* d) {                \
        ((Tr_##S*)h)->set_reference_u8(g, d);                                                        \
    }                                                                                                \
    void or_tracker_set_reference_points_##S(void* h, const uint8_t* g, const uint16_t* d, const float* xyz, \
                                             const float* nrm) {                                     \
        ((Tr_##S*)h)->set_reference_points_u8(g, d, xyz, nrm);                                       \
    }                                                                                                \
    void or_tracker_set_current_##S(void* h, const uint8_t* g, const uint16_t* d) {                  \
        ((Tr_##S*)h)->set_current_u8(g, d);                                                          \
    }                                                                                                \
    void or_tracker_track_##S(void* h) { ((Tr_##S*)h)->track(); }                                    \
    void or_tracker_fuse_current_##S(void* h) { ((Tr_##S*)h)->fuse_current(); }                      \
    void or_tracker_refresh_keyframe_##S(void* h) { ((Tr_##S*)h)->refresh_keyframe(); }              \
    void or_tracker_set_mode_##S(void* h, int mode, float lambda) {                                  \
        ((Tr_##S*)h)->prm.mode = mode; ((Tr_##S*)h)->prm.lambda = lambda;                            \
    }                                                                                                \
    void or_tracker_set_trel_##S(void* h, const double* T) {                                         \
        for (int i = 0; i < 16; i++) ((Tr_##S*)h)->Trel[i] = (R)T[i];                                \
    }                                                                                                \
    void or_tracker_get_##S(void* h, double* Trel, double* T, double* Tinc, double* global) {        \
        Tr_##S* t = (Tr_##S*)h;                                                                      \
        for (int i = 0; i < 16; i++) {                                                               \
            Trel[i] = t->Trel[i]; T[i] = t->T[i]; Tinc[i] = t->Tinc[i]; global[i] = t->global[i];    \
        }                                                                                            \
    }                                                                                                \
    int or_tracker_selected_##S(void* h, int* out) {                                                 \
        Tr_##S* t = (Tr_##S*)h;                                                                      \
        for (size_t i = 0; i < t->sel.size(); i++) out[i] = t->sel[i];                               \
        return (int)t->sel.size();                                                                   \
    }                                                                                                \
    void or_tracker_prepare_##S(void* h) {                                                           \
        Tr_##S* t = (Tr_##S*)h;                                                                      \
        if (!t->ref_prepared) t->prepare_photo_reference();                                          \
        t->prepare_current_photo();                                                                  \
        t->pending_pose_reset = false;                                                               \
    }                                                                                                \
    void or_tracker_photo_system_##S(void* h, int layer, double* A, double* b) {                     \
        ((Tr_##S*)h)->photo_system(layer, A, b);                                                     \
    }                                                                                                \
    int or_tracker_icp_system_##S(void* h, int layer, int it, double* A, double* b) {                \
        Tr_##S* t = (Tr_##S*)h;                                                                      \
        if (!t->icp_prepared) t->prepare_icp_reference();                                            \
        if (it == 0) t->prepare_current_icp();                                                       \
        t->pending_pose_reset = false;                                                               \
        return t->icp_system(layer, it, A, b) ? t->nsel_icp[layer] : 0;                              \
    }                                                                                                \
    void or_tracker_keyframe_##S(void* h, R* z, R* gray, R* w) {                                     \
        Tr_##S* t = (Tr_##S*)h;                                                                      \
        std::memcpy(z, t->kf_z.data(), sizeof(R) * t->kf_z.size());                                  \
        std::memcpy(gray, t->kf_gray0.data(), sizeof(R) * t->kf_gray0.size());                       \
        std::memcpy(w, t->kf_fw.data(), sizeof(R) * t->kf_fw.size());                                \
    }
OR_BOTH(DEF_TRACKER)

// learned pose proposal (cnn.hpp)
size_t or_cnn_num_params(const int* ch, int H0, int W0) { return cnn::Net(ch, H0, W0).nparams; }
double or_cnn_forward_backward(const int* ch, int H0, int W0, int B, const float* params, const float* x,
                               const float* y, const float* wn, int round, double* out, double* grad) {
    cnn::Net net(ch, H0, W0);
    return cnn::forward_backward(net, B, params, x, y, wn, round != 0, out, grad);
}
void or_bf16_round(const float* x, int n, float* out) {
    for (int i = 0; i < n; i++) out[i] = cnn::bf16_round(x[i]);
}

// TSDF fusion and ray casting (tsdf.hpp). prm: res, w, h; box: origin[3], dim[3], min, max
static tsdf::Grid tsdf_grid(const int* prm, const float* box) {
    tsdf::Grid g;
    tsdf::make_grid(g, prm[0], box, box + 3, prm[1], prm[2], box[6], box[7]);
    return g;
}
void or_tsdf_dist_rgbd(const int* prm, const float* box, const float* rgbd, float* dist) {
    tsdf::dist_rgbd(tsdf_grid(prm, box), rgbd, dist);
}
void or_tsdf_update(const int* prm, const float* box, float* grid, const float* dist, const float* T, const float* K) {
    tsdf::update(tsdf_grid(prm, box), grid, dist, T, K);
}
void or_tsdf_resample(const int* prm, const float* box, float* dst, const float* src, const float* T) {
    tsdf::resample(tsdf_grid(prm, box), dst, src, T);
}
void or_tsdf_raycast(const int* prm, const float* box, const float* grid, const float* T, const float* K, int refine,
                     float* xyz, float* normals, float* shade) {
    const tsdf::Grid g = tsdf_grid(prm, box);
    tsdf::raycast(g, grid, T, K, refine != 0, xyz);
    tsdf::shade(g, xyz, shade, normals);
}

}  // extern "C"
