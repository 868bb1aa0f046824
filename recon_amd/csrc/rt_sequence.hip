// Sequence driver with learned warm starts, the plain tracker as fallback and
// arbiter, and batched relocalisation (include/rt_sequence.h). Host code over
// the library's own C ABI: the single-pair trackers (rt_tracker.h), the
// batched tracker and the pose network (rt_cnn.h) and the keyframe search
// (rt_trajectory.h).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <vector>

#include "../../include/rt_cnn.h"
#include "../../include/rt_sequence.h"
#include "../../include/rt_trajectory.h"
#include "rt_host.hpp"

using rt::set_error;

namespace {

struct Mat4 {
    double m[16];
};
Mat4 ident4() {
    Mat4 r;
    for (int i = 0; i < 16; i++) r.m[i] = (i % 5 == 0) ? 1.0 : 0.0;
    return r;
}
Mat4 mul4(const Mat4& a, const Mat4& b) {
    Mat4 r;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a.m[i * 4 + k] * b.m[k * 4 + j];
            r.m[i * 4 + j] = s;
        }
    return r;
}
// inverse of a rigid transform
Mat4 inv_rigid(const Mat4& a) {
    Mat4 r;
    rt::invert_rt4(a.m, r.m);
    return r;
}
// rotation angle (degrees) and translation length (mm) of a relative pose
void angle_dist(const Mat4& T, double& deg, double& mm) {
    double c = (T.m[0] + T.m[5] + T.m[10] - 1.0) / 2.0;
    c = std::min(1.0, std::max(-1.0, c));
    deg = acos(c) * 180.0 / M_PI;
    mm = sqrt(T.m[3] * T.m[3] + T.m[7] * T.m[7] + T.m[11] * T.m[11]);
}

}  // namespace

struct rt_sequence {
    int device = 0, W = 0, H = 0;
    float calib[RT_CALIB_SIZE];
    rt_sequence_params prm;
    rt_cnn* net = nullptr;
    rt_tracker* tr = nullptr;   // the sequence's tracker
    rt_tracker* tr2 = nullptr;  // re-tracks relocalisation candidates
    rt_batch* bt = nullptr;     // batched relocalisation
    uint16_t* x = nullptr;      // network input of one pair (device)
    float* o = nullptr;         // network output of one pair (device)
    std::vector<float> grid;    // keyframe search test object
    bool relocalise = false;

    ~rt_sequence() {
        if (bt) rt_batch_destroy(bt);
        if (tr2) rt_destroy(tr2);
        if (tr) rt_destroy(tr);
        if (x) (void)hipFree(x);
        if (o) (void)hipFree(o);
    }
};

extern "C" void rt_sequence_default_params(rt_sequence_params* p) {
    if (!p) return;
    p->mode = RT_BIOBJECTIVE;
    p->lambda = 0.5f;
    p->n_select = 61440;
    p->fuse_frames = 5;
    p->max_interval = 10;
    p->dist_mm = 60.0f;
    p->angle_deg = 5.0f;
    p->relocalise = 1;
    p->reloc_keyframes = 64;
    p->reloc_sigma = 3.0f;
    p->reloc_window = 30;
    p->reloc_points = 8192;
    p->reloc_verify = 3;
    p->reloc_gate_deg = 20.0f;
    p->reloc_gate_mm = 300.0f;
    p->reloc_search_mm = 1000.0f;
    p->reloc_search_deg = 45.0f;
    p->reloc_history = 1024;
}

extern "C" int rt_sequence_create(int device, int W, int H, const float calib[RT_CALIB_SIZE], rt_cnn* net,
                                  const rt_sequence_params* params, rt_sequence** out) {
    if (!out || !calib) return set_error(RT_ERR_ARG, "rt_sequence_create: null argument");
    *out = nullptr;
    rt_sequence_params p;
    if (params) p = *params; else rt_sequence_default_params(&p);
    if (p.fuse_frames < 0 || p.max_interval < 1 || p.reloc_window < 1 || p.reloc_keyframes < 1 || p.reloc_verify < 1 ||
        p.reloc_history < 1)
        return set_error(RT_ERR_ARG, "rt_sequence_create: bad parameters");
    rt_sequence* s = new rt_sequence();
    s->device = device; s->W = W; s->H = H; s->prm = p; s->net = net;
    memcpy(s->calib, calib, sizeof(s->calib));
    s->relocalise = p.relocalise && net;
    rt_params tp;
    rt_default_params(&tp);
    tp.mode = p.mode; tp.lambda = p.lambda; tp.n_select = p.n_select;
    int rc = rt_create(device, W, H, calib, &tp, &s->tr);
    if (rc == RT_OK && s->relocalise) {
        rc = rt_create(device, W, H, calib, &tp, &s->tr2);
        if (rc == RT_OK) {
            rt_params bp;
            rt_default_params(&bp);
            bp.mode = RT_PHOTOMETRIC; bp.n_select = p.reloc_points;
            rc = rt_batch_create(device, p.reloc_keyframes, W, H, &bp, &s->bt);
        }
    }
    if (rc == RT_OK && net) {
        rt_cnn_config c;
        rt_cnn_default_config(&c);
        const size_t elems = (size_t)c.in_width * c.in_height * c.channels[0];
        if (hipMalloc((void**)&s->x, elems * 2) != hipSuccess || hipMalloc((void**)&s->o, 6 * sizeof(float)) != hipSuccess)
            rc = set_error(RT_ERR_NOMEM, "rt_sequence_create: network buffers");
    }
    if (rc == RT_OK) {
        s->grid.resize(3 * 30 * 30 * 5);
        rc = rt_grid_object(30, 30, 5, 1000.0f, 2000.0f, 3000.0f, 60.0f, s->grid.data());
    }
    if (rc != RT_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return RT_OK;
}

extern "C" void rt_sequence_destroy(rt_sequence* s) { delete s; }

namespace {

struct Ctx {
    rt_sequence* s;
    const uint8_t* g;
    const uint16_t* d;
    size_t npx;
    const uint8_t* gray(int k) const { return g + (size_t)k * npx; }
    const uint16_t* depth(int k) const { return d + (size_t)k * npx; }
};

// The network's proposal of frame k's camera pose relative to keyframe kf.
int propose(const Ctx& c, int kf, int k, Mat4& T) {
    rt_sequence* s = c.s;
    rt_pair p;
    memset(&p, 0, sizeof(p));
    p.ref_gray = c.gray(kf); p.ref_depth = c.depth(kf);
    p.cur_gray = c.gray(k); p.cur_depth = c.depth(k);
    p.width = s->W; p.height = s->H;
    int rc = rt_cnn_make_inputs(s->net, 1, &p, s->x);
    if (rc == RT_OK) rc = rt_cnn_forward(s->net, 1, s->x, s->o);
    if (rc != RT_OK) return rc;
    float y[6];
    if (hipStreamSynchronize((hipStream_t)rt_cnn_stream(s->net)) != hipSuccess ||
        hipMemcpy(y, s->o, sizeof(y), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(RT_ERR_HIP, "rt_sequence_run: network output");
    rt_cnn_config cfg;
    rt_cnn_default_config(&cfg);
    Mat4 Trel;
    if ((rc = rt_cnn_target_to_pose(&cfg, 1, y, Trel.m))) return rc;
    T = inv_rigid(Trel);
    return RT_OK;
}

int set_keyframe(const Ctx& c, rt_tracker* t, int k) {
    const int rc = rt_set_reference(t, c.gray(k), c.depth(k), 1, c.s->prm.mode, c.s->prm.lambda, 1);
    return rc < 0 ? rc : RT_OK;
}

// Track frame k on tracker t from T_init (the camera pose relative to the
// keyframe; nullptr: the previous estimate): the new relative pose and RMS.
int track(const Ctx& c, rt_tracker* t, int k, const Mat4* T_init, Mat4& T, float& rms) {
    int rc = rt_set_current(t, c.gray(k), c.depth(k), 1);
    if (rc < 0) return rc;
    rc = T_init ? rt_track_from(t, T_init->m) : rt_track(t);
    if (rc < 0) return rc;
    float Tf[16];
    if ((rc = rt_get_relative_pose(t, Tf)) || (rc = rt_get_residual_rms(t, &rms))) return rc;
    for (int i = 0; i < 16; i++) T.m[i] = Tf[i];
    return RT_OK;
}

struct Keyframe {
    int frame;
    Mat4 G;
};

// Keyframes to relocalise against: those the keyframe search finds near
// G_est, most similar first, then the most recent others.
std::vector<int> candidates(const rt_sequence* s, const std::deque<Keyframe>& kfs, const Mat4& G_est) {
    const int nk = (int)kfs.size();
    std::vector<float> poses(16 * (size_t)nk);
    for (int i = 0; i < nk; i++)
        for (int j = 0; j < 16; j++) poses[16 * (size_t)i + j] = (float)kfs[i].G.m[j];
    float Tf[16];
    for (int j = 0; j < 16; j++) Tf[j] = (float)G_est.m[j];
    std::vector<int> order(std::max(1, nk));
    std::vector<float> w(std::max(1, nk));
    int found = 0;
    std::vector<int> chosen;
    if (rt_keyframes_similar(Tf, nk, poses.data(), s->prm.reloc_search_mm, s->prm.reloc_search_deg,
                             s->calib + RT_KR_OFFSET, s->grid.data(), (int)s->grid.size() / 3, order.data(), w.data(),
                             &found) == RT_OK)
        for (int i = 0; i < found && (int)chosen.size() < s->prm.reloc_keyframes; i++) chosen.push_back(order[i]);
    for (int i = nk - 1; i >= 0 && (int)chosen.size() < s->prm.reloc_keyframes; i--)
        if (std::find(chosen.begin(), chosen.end(), i) == chosen.end()) chosen.push_back(i);
    return chosen;
}

// Frame k against the candidate keyframes in one batched launch from the
// network's proposals; the best reloc_verify re-tracked by the second
// tracker. found: the best candidate that agrees with G_est.
int relocalise(const Ctx& c, int k, const std::deque<Keyframe>& kfs, const Mat4& G_est, bool& found, Keyframe& best_kf,
               Mat4& best_T, float& best_rms) {
    rt_sequence* s = c.s;
    found = false;
    const std::vector<int> cand = candidates(s, kfs, G_est);
    const int n = (int)cand.size();
    std::vector<rt_pair> pairs(n);
    for (int i = 0; i < n; i++) {
        rt_pair& p = pairs[i];
        memset(&p, 0, sizeof(p));
        p.ref_gray = c.gray(kfs[cand[i]].frame); p.ref_depth = c.depth(kfs[cand[i]].frame);
        p.cur_gray = c.gray(k); p.cur_depth = c.depth(k);
        p.width = s->W; p.height = s->H; p.calib = s->calib;
    }
    std::vector<rt_pair_result> res(n);
    int rc = rt_batch_prepare(s->bt, n, pairs.data());
    if (rc == RT_OK) rc = rt_batch_propose(s->bt, s->net);
    if (rc == RT_OK) rc = rt_batch_retrack(s->bt, nullptr, res.data());
    if (rc != RT_OK) return rc;
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    auto rms_of = [&](int i) { return res[i].status == 0 ? (double)res[i].rms : INFINITY; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return rms_of(a) < rms_of(b); });
    for (int q = 0; q < std::min(n, s->prm.reloc_verify); q++) {
        const int j = idx[q];
        if (!std::isfinite(rms_of(j))) break;
        const Keyframe& kf = kfs[cand[j]];
        if ((rc = set_keyframe(c, s->tr2, kf.frame))) return rc;
        Mat4 Tinit, Tj;
        memcpy(Tinit.m, res[j].T, sizeof(Tinit.m));
        float rj = 0.0f;
        if ((rc = track(c, s->tr2, k, &Tinit, Tj, rj))) return rc;
        double a, t;
        angle_dist(mul4(inv_rigid(G_est), mul4(kf.G, Tj)), a, t);
        if (a > s->prm.reloc_gate_deg || t > s->prm.reloc_gate_mm) continue;
        if (!found || rj < best_rms) {
            found = true;
            best_kf = kf; best_T = Tj; best_rms = rj;
        }
    }
    return RT_OK;
}

}  // namespace

extern "C" int rt_sequence_run(rt_sequence* s, const uint8_t* gray, const uint16_t* depth, int n, double* poses_out,
                               int* trace_out, float* rms_out, int* stats) {
    if (!s || !gray || !depth || !poses_out || n < 1) return set_error(RT_ERR_ARG, "rt_sequence_run: bad argument");
    if (hipSetDevice(s->device) != hipSuccess) return set_error(RT_ERR_HIP, "hipSetDevice");
    const rt_sequence_params& P = s->prm;
    const Ctx c{s, gray, depth, (size_t)s->W * s->H};
    const Mat4 I = ident4();
    memcpy(poses_out, I.m, sizeof(I.m));
    if (trace_out) { trace_out[0] = 0; trace_out[1] = 0; }
    if (rms_out) rms_out[0] = 0.0f;
    int rc = set_keyframe(c, s->tr, 0);
    if (rc) return rc;
    int kf = 0;
    Mat4 G_kf = I, T_last = I;
    std::deque<Keyframe> keyframes{Keyframe{0, I}};
    std::deque<double> hist;
    int n_kf = 1, n_reloc = 0;
    auto spiking = [&](double rms) {
        if ((int)hist.size() < 10) return false;
        double mean = 0.0, var = 0.0;
        for (double h : hist) mean += h;
        mean /= (double)hist.size();
        for (double h : hist) var += (h - mean) * (h - mean);
        return rms > mean + (double)P.reloc_sigma * sqrt(var / (double)hist.size());
    };
    for (int k = 1; k < n; k++) {
        Mat4 T, T_init;
        float rms = 0.0f;
        if (s->net) {
            if ((rc = propose(c, kf, k, T_init))) return rc;
            if ((rc = track(c, s->tr, k, &T_init, T, rms))) return rc;
        } else if ((rc = track(c, s->tr, k, nullptr, T, rms))) {
            return rc;
        }
        bool spike = spiking(rms);
        int how = 0;  // 0 first estimate, 1 plain fallback, 2 relocalised
        if (spike && s->net) {
            Mat4 T2;
            float rms2 = 0.0f;
            if ((rc = track(c, s->tr, k, &T_last, T2, rms2))) return rc;
            if (rms2 < rms) { T = T2; rms = rms2; how = 1; }
            spike = spiking(rms);
        }
        if (spike && s->relocalise) {
            bool found = false;
            Keyframe bk{0, I};
            Mat4 bT = I;
            float brms = 0.0f;
            if ((rc = relocalise(c, k, keyframes, mul4(G_kf, T), found, bk, bT, brms))) return rc;
            if (found && brms < rms) {
                if ((rc = set_keyframe(c, s->tr, bk.frame))) return rc;
                kf = bk.frame; G_kf = bk.G;
                if ((rc = track(c, s->tr, k, &bT, T, rms))) return rc;
                n_reloc++;
                how = 2;
            }
        }
        if (how && (rc = rt_reset_to_transform(s->tr, T.m))) return rc;  // the chosen estimate
        if (trace_out) { trace_out[2 * k] = kf; trace_out[2 * k + 1] = how; }
        if (rms_out) rms_out[k] = rms;
        hist.push_back(rms);
        if ((int)hist.size() > P.reloc_window) hist.pop_front();
        const Mat4 G = mul4(G_kf, T);
        memcpy(poses_out + 16 * (size_t)k, G.m, sizeof(G.m));
        T_last = T;
        const int j = k - kf;
        if (j >= 1 && j <= P.fuse_frames) {
            if ((rc = rt_fuse_current(s->tr))) return rc;
            if (j == P.fuse_frames && (rc = rt_refresh_keyframe(s->tr))) return rc;
        }
        double a, t;
        angle_dist(T, a, t);
        if (j >= P.max_interval || t > P.dist_mm || a > P.angle_deg) {
            if ((rc = set_keyframe(c, s->tr, k))) return rc;
            kf = k; G_kf = G;
            keyframes.push_back(Keyframe{k, G});
            if ((int)keyframes.size() > P.reloc_history) keyframes.pop_front();
            n_kf++;
            T_last = I;
        }
    }
    if (stats) { stats[0] = n_kf; stats[1] = n_reloc; }
    return RT_OK;
}
