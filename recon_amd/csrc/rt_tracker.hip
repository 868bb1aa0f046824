// C ABI of the single-pair tracker: the PhotoTracker / ShapeTracker host logic
// (tracker/tracker/phototracker.cpp, shapetracker.cpp) driving the kernels of
// rt_kernels.hip on one HIP stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <string>
#include <vector>
#include "../../include/rt_tracker.h"
#include "rt_kernels.hpp"
#include "rt_host.hpp"

using namespace rt;

static thread_local std::string g_err;
namespace rt {
int set_error(int code, const char* msg) {
    g_err = msg;
    return code;
}
}  // namespace rt
extern "C" const char* rt_last_error(void) { return g_err.c_str(); }

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) {                                                                         \
            char b_[256];                                                                               \
            snprintf(b_, sizeof(b_), "%s failed: %s", #expr, hipGetErrorString(e_));                    \
            return set_error(RT_ERR_HIP, b_);                                                           \
        }                                                                                               \
    } while (0)

struct rt_tracker {
    int device = 0;
    int W = 0, H = 0, L = 3, N = 8192;
    rt_params prm;
    float calib[RT_CALIB_SIZE];
    float calib_pending[RT_CALIB_SIZE];
    bool calib_update = false;
    hipStream_t stream = nullptr;
    Dev d;
    std::vector<void*> allocs;
    uint8_t* g8[2] = {nullptr, nullptr};
    uint16_t* d16[2] = {nullptr, nullptr};
    float* rgbd = nullptr;
    float* trel_host = nullptr;   // pinned
    double* seq_poses = nullptr;  // device, grows on demand
    int seq_cap = 0;
    // PhotoTracker / ShapeTracker state
    int mode = RT_GEOMETRIC;
    float lambda = 0.9f;
    bool referenceExists = false, currentExists = false, referencePrepared = false;
    bool icpReferenceExists = false, icpPrepared = false;
    bool filteredReference = false;  // ICP keyframe from ray-cast points (setFilteredReference)
    bool icpFromPoints = false;      // the current ICP keyframe came from point maps
    bool pendingPoseReset = false;
    int refFrameNumber = 0, currentFrame = 0, referenceUpdates = 0;
    float T[16];
    double Tinc[16], global[16];
    // the Gauss-Newton launches of one frame, captured once per configuration
    // (mode, levels, iterations, solver settings) and replayed
    hipGraphExec_t est_graph = nullptr;
    int est_key[10] = {-1};
};

static void ident(float* M) { for (int i = 0; i < 16; i++) M[i] = (i % 5 == 0) ? 1.0f : 0.0f; }
static void ident(double* M) { for (int i = 0; i < 16; i++) M[i] = (i % 5 == 0) ? 1.0 : 0.0; }

extern "C" void rt_default_params(rt_params* p) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->mode = RT_GEOMETRIC;
    p->lambda = 0.9f;
    p->levels = 0;
    p->iters[0] = 1; p->iters[1] = 3; p->iters[2] = 10; p->iters[3] = 5;
    p->n_select = 8192;
    p->mad_bins = 64;
    p->robust = RT_ROBUST_TUKEY;
    p->solver = RT_SOLVER_CHOLESKY;
    p->pose_dist_threshold = 125.0f;
    p->pose_angle_threshold = 12.5f;
    p->incremental = 0;
    p->reference_updating = 0;
}

extern "C" int rt_calib_pinhole(float fx, float fy, float cx, float cy, float zmin, float zmax, float* out) {
    if (!out || !(fx > 0) || !(fy > 0) || !(zmax > zmin) || !(zmin >= 0)) return set_error(RT_ERR_ARG, "rt_calib_pinhole: bad intrinsics");
    return calib_pinhole(fx, fy, cx, cy, zmin, zmax, out);
}

static int check_params(const rt_params* p, int W, int H, int* Lout) {
    int L = p->levels > 0 ? p->levels : (W == 640 ? 4 : 3);
    if (L < 1 || L > 4) return set_error(RT_ERR_ARG, "levels must be in 1..4");
    int m = 1 << (L - 1);
    if (m < 4) m = 4;
    if (W % (2 * m) != 0 || H % m != 0 || W < 32 || H < 32) return set_error(RT_ERR_ARG, "width/height must be multiples of 4 and 2^(levels-1) (width also of 2^levels)");
    if ((W * H) / 512 > 1024) return set_error(RT_ERR_ARG, "width*height/512 must be at most 1024");
    if (p->n_select < 256 || p->n_select % 256 != 0 || p->n_select > W * H / 2) return set_error(RT_ERR_ARG, "n_select must be a multiple of 256 in [256, W*H/2]");
    if (p->mad_bins != 64 && p->mad_bins != 256) return set_error(RT_ERR_ARG, "mad_bins must be 64 or 256");
    if (p->robust != RT_ROBUST_TUKEY && p->robust != RT_ROBUST_HUBER) return set_error(RT_ERR_ARG, "bad robust kind");
    if (p->solver != RT_SOLVER_CHOLESKY && p->solver != RT_SOLVER_CG) return set_error(RT_ERR_ARG, "bad solver");
    if (p->mode < 0 || p->mode > 2) return set_error(RT_ERR_ARG, "bad mode");
    for (int i = 0; i < 4; i++) if (p->iters[i] < 0 || p->iters[i] > 1000) return set_error(RT_ERR_ARG, "bad iteration count");
    *Lout = L;
    return RT_OK;
}

template <class T>
static int dalloc(rt_tracker* t, T** p, size_t count) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e != hipSuccess) return set_error(RT_ERR_NOMEM, "hipMalloc failed");
    hipMemset(q, 0, count * sizeof(T) + 16);
    t->allocs.push_back(q);
    *p = (T*)q;
    return RT_OK;
}

static void free_all(rt_tracker* t) {
    if (t->stream) hipStreamSynchronize(t->stream);
    if (t->est_graph) hipGraphExecDestroy(t->est_graph);
    t->est_graph = nullptr;
    for (void* q : t->allocs) hipFree(q);
    t->allocs.clear();
    if (t->trel_host) hipHostFree(t->trel_host);
    t->trel_host = nullptr;
    if (t->stream) hipStreamDestroy(t->stream);
    t->stream = nullptr;
}

extern "C" int rt_create(int device, int W, int H, const float* calib, const rt_params* params, rt_tracker** out) {
    if (!out || !calib) return set_error(RT_ERR_ARG, "rt_create: null argument");
    *out = nullptr;
    rt_params p;
    if (params) p = *params; else rt_default_params(&p);
    int L = 0;
    int rc = check_params(&p, W, H, &L);
    if (rc) return rc;
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return set_error(RT_ERR_ARG, "rt_create: no such device");
    HIP_TRY(hipSetDevice(device));
    rt_tracker* t = new rt_tracker();
    t->device = device; t->W = W; t->H = H; t->L = L; t->N = p.n_select; t->prm = p;
    t->mode = p.mode; t->lambda = p.lambda < 0 ? 0.0f : (p.lambda > 1 ? 1.0f : p.lambda);
    memcpy(t->calib, calib, sizeof(t->calib));
    ident(t->T); ident(t->Tinc); ident(t->global);
    if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) { delete t; return set_error(RT_ERR_HIP, "stream"); }
    Dev& d = t->d;
    memset(&d, 0, sizeof(d));
    d.W = W; d.H = H; d.L = L; d.N = p.n_select;
    size_t npx = (size_t)W * H;
    int G1024 = (int)((npx + 1023) / 1024);
    bool ok = true;
#define A(ptr, n) ok = ok && dalloc(t, &(ptr), (size_t)(n)) == RT_OK
    float* cal = nullptr;
    A(cal, RT_CALIB_SIZE);
    d.calib = cal;
    for (int l = 0; l < 4; l++) {
        size_t nl = (size_t)(W >> l) * (H >> l);
        if (l < L) { A(d.ref.gray[l], nl); A(d.cur.gray[l], nl); }
    }
    A(d.ref.z, npx); A(d.cur.z, npx); A(d.kf_w, npx);
    A(d.gmag, npx); A(d.ghist, 256); A(d.seg, 2600); A(d.sel, p.n_select); A(d.pts, p.n_select);
    for (int l = 0; l < L; l++) A(d.jac[l], (jac_bytes(p.n_select) + 15) / 16);
    A(d.res, p.n_select); A(d.wts, p.n_select); A(d.madhist, 512);
    A(d.part, 27 * PART_STRIDE); A(d.part2, 27 * PART_STRIDE); A(d.part3, 27 * RAW_STRIDE); A(d.sys, 64); A(d.trel, 16);
    for (int l = 0; l < L; l++) {
        size_t nl = (size_t)(W >> l) * (H >> l);
        A(d.rpos[l], nl); A(d.rnrm[l], nl); A(d.cpos[l], nl); A(d.icp_sel[l], nl); A(d.icp_r2[l], nl);
    }
    A(d.rfull, npx); A(d.icp_cnt, G1024); A(d.icp_bits, (size_t)G1024 * 16); A(d.icp_var, 16); A(d.icp_n, 4);
    A(d.bsum, 3 * (size_t)G1024); A(d.pose, 32);
    A(d.dbg_err, npx); A(d.dbg_ref, npx);
    for (int l = 0; l < L; l++) {
        size_t nl = (size_t)(W >> l) * (H >> l);
        A(d.dbg_icp[l], nl); A(d.dbg_icp_ref[l], nl);
    }
    for (int k = 0; k < 2; k++) { A(t->g8[k], npx); A(t->d16[k], npx); }
    A(t->rgbd, 4 * npx);
#undef A
    if (!ok || hipHostMalloc((void**)&t->trel_host, 64 * sizeof(float), 0) != hipSuccess) {
        free_all(t); delete t;
        return set_error(RT_ERR_NOMEM, "rt_create: allocation failed");
    }
    if (hipMemcpy(cal, calib, RT_CALIB_SIZE * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
        free_all(t); delete t;
        return set_error(RT_ERR_HIP, "rt_create: calibration upload failed");
    }
    float I[16]; ident(I);
    double pose0[32]; ident(pose0); ident(pose0 + 16);
    hipMemcpy(d.trel, I, sizeof(I), hipMemcpyHostToDevice);
    hipMemcpy(d.pose, pose0, sizeof(pose0), hipMemcpyHostToDevice);
    hipDeviceSynchronize();
    *out = t;
    return RT_OK;
}

extern "C" void rt_destroy(rt_tracker* t) {
    if (!t) return;
    hipSetDevice(t->device);
    free_all(t);
    delete t;
}

static SolveCfg solve_cfg(const rt_tracker* t) {
    SolveCfg c;
    c.solver = t->prm.solver;
    c.mode = t->mode;
    c.lambda = t->lambda;
    c.scale_out = 1e3f;
    c.robust = t->prm.robust;
    c.mad_bins = t->prm.mad_bins;
    c.scale_in = 1e-3f;
    return c;
}

static int upload_frame(rt_tracker* t, int slot, const uint8_t* g, const uint16_t* dd, int on_device,
                        const uint8_t** g_dev, const uint16_t** d_dev) {
    if (!g || !dd) return set_error(RT_ERR_ARG, "null image");
    size_t npx = (size_t)t->W * t->H;
    if (on_device) { *g_dev = g; *d_dev = dd; return RT_OK; }
    HIP_TRY(hipMemcpyAsync(t->g8[slot], g, npx, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipMemcpyAsync(t->d16[slot], dd, npx * 2, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));  // host frames are read during the call only
    *g_dev = t->g8[slot]; *d_dev = t->d16[slot];
    return RT_OK;
}

extern "C" int rt_set_current(rt_tracker* t, const uint8_t* gray, const uint16_t* depth, int on_device) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    const uint8_t* g; const uint16_t* dd;
    int rc = upload_frame(t, 1, gray, depth, on_device, &g, &dd);
    if (rc) return rc;
    launch_convert_pyramid(t->d, t->d.cur, g, dd, nullptr, true, false, t->stream);
    HIP_TRY(hipGetLastError());
    t->currentExists = true;
    return RT_OK;
}

extern "C" int rt_set_current_rgbd(rt_tracker* t, const float* rgbd, int on_device) {
    if (!t || !rgbd) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    const float* src = rgbd;
    if (!on_device) {
        HIP_TRY(hipMemcpyAsync(t->rgbd, rgbd, (size_t)t->W * t->H * 16, hipMemcpyHostToDevice, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));  // host frames are read during the call only
        src = t->rgbd;
    }
    launch_convert_pyramid(t->d, t->d.cur, nullptr, nullptr, src, true, false, t->stream);
    HIP_TRY(hipGetLastError());
    t->currentExists = true;
    return RT_OK;
}

// referenceOutdated (phototracker.cpp:292-296, shapetracker.cpp:188-193)
static bool photo_outdated(const rt_tracker* t) {
    if (!t->referenceExists) return true;
    if (t->prm.incremental) return true;
    return t->prm.reference_updating &&
           pose_difference_exceeded(t->T, t->prm.pose_dist_threshold, t->prm.pose_angle_threshold, false);
}
static bool icp_outdated(const rt_tracker* t) {
    if (!t->icpReferenceExists) return true;
    if (t->prm.incremental) return true;
    return t->prm.reference_updating &&
           pose_difference_exceeded(t->T, t->prm.pose_dist_threshold, t->prm.pose_angle_threshold, true);
}

static void set_mode(rt_tracker* t, int mode) { t->mode = mode; }

// updatePhotometricReference / ShapeTracker::setReferencePoints on a converted
// keyframe: the ICP cloud from the measured depth, or from ray-cast points and
// normals when the filtered reference is on (shapetracker.cpp:199-207)
static int update_reference(rt_tracker* t, const uint8_t* g, const uint16_t* dd, const float* rgbd, const float* xyz,
                            const float* nrm, bool photo, bool icp) {
    launch_convert_pyramid(t->d, t->d.ref, g, dd, rgbd, false, true, t->stream);
    HIP_TRY(hipGetLastError());
    // fusion weights restart at one
    launch_fill(t->d.kf_w, 1.0f, t->W * t->H, t->stream);
    if (photo) { t->referenceExists = true; t->referencePrepared = false; }
    if (icp) {
        if (t->filteredReference) launch_ref_oriented(t->d, xyz, nrm, t->stream);
        else launch_ref_cloud(t->d, t->stream);
        HIP_TRY(hipGetLastError());
        t->icpFromPoints = t->filteredReference;
        t->icpReferenceExists = true; t->icpPrepared = false;
    }
    t->pendingPoseReset = true;
    t->refFrameNumber = t->currentFrame;
    t->referenceUpdates++;
    return RT_OK;
}

static int set_reference_impl(rt_tracker* t, const uint8_t* g, const uint16_t* dd, const float* rgbd, const float* xyz,
                              const float* nrm, int mode, float lambda, int force) {
    if (mode < 0 || mode > 2) return set_error(RT_ERR_ARG, "bad mode");
    // phototracker.cpp:313: a filtered reference needs the ray-cast maps
    if (t->filteredReference && (!xyz || !nrm))
        return set_error(RT_ERR_ARG, "filtered reference: ray-cast points and normals are required");
    int rc;
    if (mode == RT_PHOTOMETRIC) {
        if (force || photo_outdated(t)) {
            set_mode(t, RT_PHOTOMETRIC);
            if ((rc = update_reference(t, g, dd, rgbd, xyz, nrm, true, false))) return rc;
            return 1;
        }
    } else if (mode == RT_GEOMETRIC) {
        if (force || icp_outdated(t)) {
            set_mode(t, RT_GEOMETRIC);
            if ((rc = update_reference(t, g, dd, rgbd, xyz, nrm, false, true))) return rc;
            return 1;
        }
    } else {
        if (force || photo_outdated(t) || icp_outdated(t)) {
            set_mode(t, RT_BIOBJECTIVE);
            t->lambda = lambda < 0 ? 0.0f : (lambda > 1 ? 1.0f : lambda);
            if ((rc = update_reference(t, g, dd, rgbd, xyz, nrm, true, true))) return rc;
            return 1;
        }
    }
    return 0;
}

extern "C" int rt_set_reference(rt_tracker* t, const uint8_t* gray, const uint16_t* depth, int on_device,
                                int suggested_mode, float lambda, int force) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    const uint8_t* g; const uint16_t* dd;
    int rc = upload_frame(t, 0, gray, depth, on_device, &g, &dd);
    if (rc) return rc;
    return set_reference_impl(t, g, dd, nullptr, nullptr, nullptr, suggested_mode, lambda, force);
}

extern "C" int rt_set_reference_rgbd(rt_tracker* t, const float* rgbd, int on_device, int suggested_mode,
                                     float lambda, int force) {
    if (!t || !rgbd) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    const float* src = rgbd;
    if (!on_device) {
        HIP_TRY(hipMemcpyAsync(t->rgbd, rgbd, (size_t)t->W * t->H * 16, hipMemcpyHostToDevice, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));  // host frames are read during the call only
        src = t->rgbd;
    }
    return set_reference_impl(t, nullptr, nullptr, src, nullptr, nullptr, suggested_mode, lambda, force);
}

extern "C" int rt_set_reference_points(rt_tracker* t, const float* xyz, const float* normals, const uint8_t* gray,
                                       const uint16_t* depth, int on_device, int suggested_mode, float lambda,
                                       int force) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    const uint8_t* g; const uint16_t* dd;
    int rc = upload_frame(t, 0, gray, depth, on_device, &g, &dd);
    if (rc) return rc;
    return set_reference_impl(t, g, dd, nullptr, xyz, normals, suggested_mode, lambda, force);
}

extern "C" int rt_set_reference_points_rgbd(rt_tracker* t, const float* xyz, const float* normals, const float* rgbd,
                                            int on_device, int suggested_mode, float lambda, int force) {
    if (!t || !rgbd) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    const float* src = rgbd;
    if (!on_device) {
        HIP_TRY(hipMemcpyAsync(t->rgbd, rgbd, (size_t)t->W * t->H * 16, hipMemcpyHostToDevice, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));  // host frames are read during the call only
        src = t->rgbd;
    }
    return set_reference_impl(t, nullptr, nullptr, src, xyz, normals, suggested_mode, lambda, force);
}

// PhotoTracker::updateErrorImage (phototracker.cpp:388-403): the photometric
// error image at level 0, or the geometric ones of every level
// (ShapeTracker::updateErrorImage, shapetracker.cpp:237-239), under the
// current pose of the prepared keyframe and frame.
extern "C" int rt_update_error_image(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    if (t->mode == RT_PHOTOMETRIC) {
        if (!t->referencePrepared) return set_error(RT_ERR_STATE, "rt_update_error_image: keyframe not prepared");
        launch_error_image_photo(t->d, t->stream);
    }
    if (t->mode == RT_GEOMETRIC || t->mode == RT_BIOBJECTIVE) {
        if (!t->icpPrepared) return set_error(RT_ERR_STATE, "rt_update_error_image: keyframe not prepared");
        launch_error_image_icp(t->d, t->stream);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}
static int copy_image(rt_tracker* t, const float* src, size_t n, float* out, size_t count) {
    if (!out || count < n) return set_error(RT_ERR_ARG, "image buffer too small");
    HIP_TRY(hipSetDevice(t->device));
    HIP_TRY(hipMemcpyAsync(out, src, n * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}
// getDebugImage1C (phototracker.cpp:734-744): photometric mode the W x H error
// image whatever the level; the other modes the geometric one of `level`.
extern "C" int rt_get_debug_image(rt_tracker* t, int level, float* out, size_t count) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (t->mode == RT_PHOTOMETRIC) return copy_image(t, t->d.dbg_err, (size_t)t->W * t->H, out, count);
    if (level < 0 || level >= t->L) return set_error(RT_ERR_ARG, "bad level");
    return copy_image(t, t->d.dbg_icp[level], (size_t)(t->W >> level) * (t->H >> level), out, count);
}
// getRefImage1C (phototracker.cpp:746-755): photometric mode the keyframe's
// normalised depth (W x H); the other modes the shape tracker's reference
// image of `level`, which the reference never writes (its normalising kernel,
// xyz6ToDepthNormalizedKernel, rgb2gray.cu:125-131, stores nothing): zeros.
extern "C" int rt_get_ref_image(rt_tracker* t, int level, float* out, size_t count) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (t->mode == RT_PHOTOMETRIC) return copy_image(t, t->d.dbg_ref, (size_t)t->W * t->H, out, count);
    if (level < 0 || level >= t->L) return set_error(RT_ERR_ARG, "bad level");
    return copy_image(t, t->d.dbg_icp_ref[level], (size_t)(t->W >> level) * (t->H >> level), out, count);
}

extern "C" int rt_set_filtered_reference(rt_tracker* t, int flag) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    t->filteredReference = flag != 0;
    return RT_OK;
}
extern "C" int rt_get_filtered_reference(const rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    return t->filteredReference ? 1 : 0;
}

static bool ready(const rt_tracker* t) {
    bool photo = t->referenceExists && t->currentExists;
    bool icp = t->icpReferenceExists && t->currentExists;
    if (t->mode == RT_PHOTOMETRIC) return photo;
    if (t->mode == RT_GEOMETRIC) return icp;
    return photo && icp;
}

// gpuPreProcess (phototracker.cpp:596-636). seq: device pose state in use.
static int preprocess(rt_tracker* t, bool seq) {
    if (t->calib_update) {
        HIP_TRY(hipMemcpyAsync((void*)t->d.calib, t->calib_pending, sizeof(t->calib), hipMemcpyHostToDevice, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        memcpy(t->calib, t->calib_pending, sizeof(t->calib));
        t->calib_update = false;
    }
    int m = t->mode;
    bool reset = false;
    if ((m == RT_GEOMETRIC || m == RT_BIOBJECTIVE) && !t->icpPrepared) {
        if (t->pendingPoseReset) reset = true;
        launch_ref_cloud_pyramid(t->d, t->stream);
        t->icpPrepared = true;
    }
    if ((m == RT_PHOTOMETRIC || m == RT_BIOBJECTIVE) && !t->referencePrepared) {
        if (t->pendingPoseReset) reset = true;
        launch_prepare_photo_reference(t->d, t->stream);
        t->referencePrepared = true;
    }
    if (reset) {
        launch_reset_pose(t->d.trel, seq ? t->d.pose : nullptr, t->stream);
        ident(t->T);
        t->pendingPoseReset = false;
    }
    if (m == RT_GEOMETRIC || m == RT_BIOBJECTIVE) launch_cur_cloud(t->d, t->stream);
    HIP_TRY(hipGetLastError());
    return RT_OK;
}

// gpuEstimateIncrement (phototracker.cpp:666-689)
static void estimate_launches(rt_tracker* t, const SolveCfg& c) {
    int m = t->mode;
    for (int layer = t->L - 1; layer >= 0; layer--) {
        int n = t->prm.iters[t->L - 1 - layer];
        for (int j = 0; j < n; j++) {
            if (m == RT_PHOTOMETRIC) {
                launch_photo_iteration(t->d, layer, c, FINISH_SOLVE, t->stream);
            } else if (m == RT_GEOMETRIC) {
                if (j == 0) launch_icp_select(t->d, layer, t->stream);
                launch_icp_system(t->d, layer, c, FINISH_SOLVE, t->stream);
            } else {
                launch_biobjective_iteration(t->d, layer, c, FINISH_SOLVE, j == 0, t->stream);
            }
        }
    }
}

// gpuEstimateIncrement (phototracker.cpp:666-689): every level's iterations.
// The launches take only fixed device buffers and the solve settings, so they
// are captured into a graph once per configuration and replayed per frame
// (RT_EAGER set in the environment launches them one by one instead).
static int estimate(rt_tracker* t) {
    SolveCfg c = solve_cfg(t);
    int lam_bits;
    memcpy(&lam_bits, &c.lambda, sizeof(lam_bits));
    const int key[10] = {t->mode, t->L, t->prm.iters[0], t->prm.iters[1], t->prm.iters[2], t->prm.iters[3],
                         c.solver, c.robust, c.mad_bins, lam_bits};
    if (getenv("RT_EAGER")) {
        estimate_launches(t, c);
    } else {
        if (!t->est_graph || memcmp(key, t->est_key, sizeof(key))) {
            if (t->est_graph) { HIP_TRY(hipGraphExecDestroy(t->est_graph)); t->est_graph = nullptr; }
            hipGraph_t g;
            HIP_TRY(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
            estimate_launches(t, c);
            HIP_TRY(hipStreamEndCapture(t->stream, &g));
            hipError_t e = hipGraphInstantiate(&t->est_graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { t->est_graph = nullptr; return set_error(RT_ERR_HIP, "graph instantiate"); }
            memcpy(t->est_key, key, sizeof(key));
        }
        HIP_TRY(hipGraphLaunch(t->est_graph, t->stream));
    }
    HIP_TRY(hipGetLastError());
    t->currentFrame++;
    return RT_OK;
}

extern "C" int rt_track(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    if (!ready(t)) return 0;
    int rc = preprocess(t, false);
    if (rc) return rc;
    if ((rc = estimate(t))) return rc;
    HIP_TRY(hipMemcpyAsync(t->trel_host, t->d.trel, 16 * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    apply_increment(t->trel_host, t->T, t->Tinc, t->global);
    return 1;
}

// rt_track with the Gauss-Newton loop started from a given pose of the current
// camera relative to the keyframe (a warm start, e.g. a learned proposal)
// instead of the previous estimate. The pose bookkeeping (applyIncrement) is
// that of rt_track.
extern "C" int rt_track_from(rt_tracker* t, const double* T_init) {
    if (!t || !T_init) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    if (!ready(t)) return 0;
    int rc = preprocess(t, false);
    if (rc) return rc;
    float Tf[16], iT[16];
    for (int i = 0; i < 16; i++) Tf[i] = (float)T_init[i];
    invert_rt4f(Tf, iT);
    memcpy(t->trel_host, iT, sizeof(iT));  // pinned staging, read by the copy before anything rewrites it
    HIP_TRY(hipMemcpyAsync(t->d.trel, t->trel_host, 16 * sizeof(float), hipMemcpyHostToDevice, t->stream));
    if ((rc = estimate(t))) return rc;
    HIP_TRY(hipMemcpyAsync(t->trel_host, t->d.trel, 16 * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    apply_increment(t->trel_host, t->T, t->Tinc, t->global);
    return 1;
}

// Weighted photometric residual RMS of the last Gauss-Newton iteration,
// sqrt(sum (w r)^2 / sum w^2) over the selected points (0 when all weights vanish).
extern "C" int rt_get_residual_rms(rt_tracker* t, float* rms) {
    if (!t || !rms) return set_error(RT_ERR_ARG, "null argument");
    if (t->mode == RT_GEOMETRIC) return set_error(RT_ERR_STATE, "no photometric residuals in geometric mode");
    HIP_TRY(hipSetDevice(t->device));
    std::vector<float> rz(2 * (size_t)t->N), w(t->N);
    HIP_TRY(hipMemcpyAsync(rz.data(), t->d.res, rz.size() * 4, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipMemcpyAsync(w.data(), t->d.wts, w.size() * 4, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    double a = 0.0, b = 0.0;
    for (int i = 0; i < t->N; i++) {
        double wr = (double)w[i] * rz[2 * (size_t)i];
        a += wr * wr;
        b += (double)w[i] * w[i];
    }
    *rms = b > 0.0 ? (float)sqrt(a / b) : 0.0f;
    return RT_OK;
}

extern "C" int rt_get_global_pose(const rt_tracker* t, double* T) {
    if (!t || !T) return set_error(RT_ERR_ARG, "null argument");
    memcpy(T, t->global, sizeof(t->global));
    return RT_OK;
}
extern "C" int rt_get_relative_pose(const rt_tracker* t, float* T) {
    if (!t || !T) return set_error(RT_ERR_ARG, "null argument");
    memcpy(T, t->T, sizeof(t->T));
    return RT_OK;
}
extern "C" int rt_get_increment(const rt_tracker* t, double* T) {
    if (!t || !T) return set_error(RT_ERR_ARG, "null argument");
    memcpy(T, t->Tinc, sizeof(t->Tinc));
    return RT_OK;
}

// clearPose (phototracker.cpp:245-253)
extern "C" int rt_clear_pose(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    ident(t->global); ident(t->Tinc); ident(t->T);
    launch_reset_pose(t->d.trel, nullptr, t->stream);
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}
// resetTracking (phototracker.cpp:231-243)
extern "C" int rt_reset_tracking(rt_tracker* t, int mode) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (mode < 0 || mode > 2) return set_error(RT_ERR_ARG, "bad mode");
    int rc = rt_clear_pose(t);
    if (rc) return rc;
    t->referenceExists = t->currentExists = t->referencePrepared = false;
    t->icpReferenceExists = t->icpPrepared = false;
    t->refFrameNumber = 0; t->currentFrame = 0; t->referenceUpdates = 0;
    set_mode(t, mode);
    return RT_OK;
}
// resetToTransform (phototracker.cpp:545-561)
extern "C" int rt_reset_to_transform(rt_tracker* t, const double* T) {
    if (!t || !T) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    for (int i = 0; i < 16; i++) t->T[i] = (float)T[i];
    ident(t->Tinc);
    float iT[16];
    invert_rt4f(t->T, iT);
    HIP_TRY(hipMemcpyAsync(t->d.trel, iT, sizeof(iT), hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}
extern "C" int rt_set_mode(rt_tracker* t, int mode) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (mode < 0 || mode > 2) return set_error(RT_ERR_ARG, "bad mode");
    set_mode(t, mode);
    return RT_OK;
}
extern "C" int rt_get_mode(const rt_tracker* t) { return t ? t->mode : set_error(RT_ERR_ARG, "null tracker"); }
extern "C" int rt_set_lambda(rt_tracker* t, float lambda) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    t->lambda = lambda < 0 ? 0.0f : (lambda > 1 ? 1.0f : lambda);
    return RT_OK;
}
extern "C" float rt_get_lambda(const rt_tracker* t) { return t ? t->lambda : -1.0f; }
extern "C" int rt_set_iteration_counts(rt_tracker* t, const int* n) {
    if (!t || !n) return set_error(RT_ERR_ARG, "null argument");
    for (int i = 0; i < t->L; i++) if (n[i] < 0 || n[i] > 1000) return set_error(RT_ERR_ARG, "bad iteration count");
    for (int i = 0; i < t->L; i++) t->prm.iters[i] = n[i];
    return RT_OK;
}
extern "C" int rt_get_iteration_counts(const rt_tracker* t, int* n) {
    if (!t || !n) return set_error(RT_ERR_ARG, "null argument");
    memcpy(n, t->prm.iters, sizeof(t->prm.iters));
    return RT_OK;
}
extern "C" int rt_set_reference_updating(rt_tracker* t, int flag) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    t->prm.reference_updating = flag ? 1 : 0;
    return RT_OK;
}
extern "C" int rt_set_incremental_mode(rt_tracker* t, int flag) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    t->prm.incremental = flag ? 1 : 0;
    return RT_OK;
}
extern "C" int rt_signal_calibration_update(rt_tracker* t, const float* calib) {
    if (!t || !calib) return set_error(RT_ERR_ARG, "null argument");
    memcpy(t->calib_pending, calib, sizeof(t->calib_pending));
    t->calib_update = true;
    return RT_OK;
}
extern "C" int rt_get_layers(const rt_tracker* t) { return t ? t->L : set_error(RT_ERR_ARG, "null tracker"); }
extern "C" int rt_num_reference_updates(const rt_tracker* t) {
    return t ? t->referenceUpdates : set_error(RT_ERR_ARG, "null tracker");
}

// listSelected / listSelectedCurCuda (phototracker.cpp:405-446)
extern "C" int rt_list_selected(rt_tracker* t, float* cur_xy, float* ref_xy, float* weight, int capacity, int* n) {
    if (!t || !n) return set_error(RT_ERR_ARG, "null argument");
    *n = 0;
    if (!t->referencePrepared) return RT_OK;
    if (capacity < t->N) return set_error(RT_ERR_ARG, "capacity below the selected point count");
    HIP_TRY(hipSetDevice(t->device));
    std::vector<float4> pts(t->N);
    std::vector<float> w(t->N);
    float trel[16];
    HIP_TRY(hipMemcpyAsync(pts.data(), t->d.pts, t->N * sizeof(float4), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipMemcpyAsync(w.data(), t->d.wts, t->N * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipMemcpyAsync(trel, t->d.trel, sizeof(trel), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    const float* K = t->calib + RT_KR_OFFSET;
    const float* TLR = t->calib + RT_TLR_OFFSET;
    for (int i = 0; i < t->N; i++) {
        float p[3] = {pts[i].x, pts[i].y, pts[i].z};
        float q[3];
        for (int pass = 0; pass < 2; pass++) {
            const float* Tm = pass == 0 ? nullptr : trel;
            float a[3];
            for (int k = 0; k < 3; k++)
                a[k] = Tm ? Tm[k * 4] * p[0] + Tm[k * 4 + 1] * p[1] + Tm[k * 4 + 2] * p[2] + Tm[k * 4 + 3] : p[k];
            for (int k = 0; k < 3; k++) q[k] = TLR[k * 4] * a[0] + TLR[k * 4 + 1] * a[1] + TLR[k * 4 + 2] * a[2] + TLR[k * 4 + 3];
            float u = K[0] * (q[0] / q[2]) + K[2], v = K[4] * (q[1] / q[2]) + K[5];
            float* dst = pass == 0 ? ref_xy : cur_xy;
            if (dst) { dst[2 * i] = u; dst[2 * i + 1] = v; }
        }
        if (weight) weight[i] = w[i];
    }
    *n = t->N;
    return RT_OK;
}

extern "C" int rt_fuse_current(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (!t->referenceExists && !t->icpReferenceExists) return set_error(RT_ERR_STATE, "no keyframe");
    if (!t->currentExists) return set_error(RT_ERR_STATE, "no current frame");
    HIP_TRY(hipSetDevice(t->device));
    launch_fuse(t->d, t->stream);
    HIP_TRY(hipGetLastError());
    return RT_OK;
}

// Re-prepare the keyframe from its (fused) level-0 images, keeping the pose.
static int refresh_keyframe(rt_tracker* t) {
    launch_convert_pyramid(t->d, t->d.ref, nullptr, nullptr, nullptr, false, true, t->stream);
    if (t->icpReferenceExists && !t->icpFromPoints) {  // a cloud from ray-cast points stays as it is
        launch_ref_cloud(t->d, t->stream);
        t->icpPrepared = false;
    }
    if (t->referenceExists) t->referencePrepared = false;
    HIP_TRY(hipGetLastError());
    return RT_OK;
}
extern "C" int rt_refresh_keyframe(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    return refresh_keyframe(t);
}

extern "C" void* rt_stream(rt_tracker* t) { return t ? (void*)t->stream : nullptr; }
extern "C" int rt_synchronize(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}

extern "C" int rt_track_sequence(rt_tracker* t, const uint8_t* gray, const uint16_t* depth, int n, int interval,
                                 int fuse_frames, double* poses_out) {
    if (!t || !gray || !depth || !poses_out || n < 1) return set_error(RT_ERR_ARG, "rt_track_sequence: bad argument");
    if (interval < 1) return set_error(RT_ERR_ARG, "keyframe_interval must be >= 1");
    if (fuse_frames < 0 || fuse_frames >= interval) return set_error(RT_ERR_ARG, "fuse_frames must be in [0, interval)");
    if (t->filteredReference)
        return set_error(RT_ERR_STATE, "rt_track_sequence keys on measured frames: turn the filtered reference off");
    HIP_TRY(hipSetDevice(t->device));
    if (t->seq_cap < n) {
        if (t->seq_poses) {
            // drop it from the allocation list too: destroy must not free it again
            t->allocs.erase(std::remove(t->allocs.begin(), t->allocs.end(), (void*)t->seq_poses), t->allocs.end());
            HIP_TRY(hipFree(t->seq_poses));
        }
        t->seq_poses = nullptr;
        t->seq_cap = 0;
        HIP_TRY(hipMalloc((void**)&t->seq_poses, (size_t)n * 16 * sizeof(double)));
        t->allocs.push_back(t->seq_poses);
        t->seq_cap = n;
    }
    size_t npx = (size_t)t->W * t->H;
    double pose0[32]; ident(pose0); ident(pose0 + 16);
    HIP_TRY(hipMemcpyAsync(t->d.pose, pose0, sizeof(pose0), hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipMemcpyAsync(t->seq_poses, pose0, 16 * sizeof(double), hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));  // pose0 lives on this function's stack
    int mode = t->mode;
    int rc = set_reference_impl(t, gray, depth, nullptr, nullptr, nullptr, mode, t->lambda, 1);
    if (rc < 0) return rc;
    int kf = 0;
    for (int k = 1; k < n; k++) {
        if (k - 1 > 0 && (k - 1) % interval == 0) {
            rc = set_reference_impl(t, gray + (k - 1) * npx, depth + (k - 1) * npx, nullptr, nullptr, nullptr, mode,
                                    t->lambda, 1);
            if (rc < 0) return rc;
            kf = k - 1;
        }
        launch_convert_pyramid(t->d, t->d.cur, gray + k * npx, depth + k * npx, nullptr, true, false, t->stream);
        t->currentExists = true;
        if ((rc = preprocess(t, true))) return rc;
        if ((rc = estimate(t))) return rc;
        launch_apply_increment(t->d, t->seq_poses + (size_t)k * 16, t->stream);
        int j = k - kf;
        if (fuse_frames > 0 && j >= 1 && j <= fuse_frames) {
            launch_fuse(t->d, t->stream);
            if (j == fuse_frames && (rc = refresh_keyframe(t))) return rc;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(poses_out, t->seq_poses, (size_t)n * 16 * sizeof(double), hipMemcpyDeviceToHost, t->stream));
    double pose[32];
    HIP_TRY(hipMemcpyAsync(pose, t->d.pose, sizeof(pose), hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    for (int i = 0; i < 16; i++) { t->T[i] = (float)pose[i]; t->global[i] = pose[16 + i]; }
    return RT_OK;
}

// ------------------------------------------------------------------ inspection
extern "C" int rt_debug_read(rt_tracker* t, int what, int level, void* host, size_t bytes) {
    if (!t || !host) return set_error(RT_ERR_ARG, "null argument");
    if (level < 0 || level >= t->L) return set_error(RT_ERR_ARG, "bad level");
    HIP_TRY(hipSetDevice(t->device));
    size_t nl = (size_t)(t->W >> level) * (t->H >> level), npx = (size_t)t->W * t->H;
    const void* src = nullptr;
    size_t n = 0;
    Dev& d = t->d;
    switch (what) {
        case RT_DBG_REF_GRAY: src = d.ref.gray[level]; n = nl * 4; break;
        case RT_DBG_CUR_GRAY: src = d.cur.gray[level]; n = nl * 4; break;
        case RT_DBG_REF_Z: src = d.ref.z; n = npx * 4; break;
        case RT_DBG_CUR_Z: src = d.cur.z; n = npx * 4; break;
        case RT_DBG_GRAD_MAG: src = d.gmag; n = npx * 4; break;
        case RT_DBG_GRAD_HIST: src = d.ghist; n = 256 * 4; break;
        case RT_DBG_SELECTED: src = d.sel; n = (size_t)t->N * 4; break;
        case RT_DBG_POINTS: src = d.pts; n = (size_t)t->N * 16; break;
        case RT_DBG_JACOBIAN: src = d.jac[level]; n = jac_bytes(t->N); break;
        case RT_DBG_RESIDUAL: src = d.res; n = (size_t)t->N * 8; break;
        case RT_DBG_TREL: src = d.trel; n = 64; break;
        case RT_DBG_REF_CLOUD: {
            if (bytes < nl * 32) return set_error(RT_ERR_ARG, "buffer too small");
            HIP_TRY(hipMemcpyAsync(host, d.rpos[level], nl * 16, hipMemcpyDeviceToHost, t->stream));
            HIP_TRY(hipMemcpyAsync((char*)host + nl * 16, d.rnrm[level], nl * 16, hipMemcpyDeviceToHost, t->stream));
            HIP_TRY(hipStreamSynchronize(t->stream));
            return RT_OK;
        }
        case RT_DBG_CUR_CLOUD: src = d.cpos[level]; n = nl * 16; break;
        case RT_DBG_ICP_SELECTED: src = d.icp_sel[level]; n = nl * 4; break;
        case RT_DBG_ICP_COUNT: src = d.icp_n; n = 16; break;
        case RT_DBG_ICP_R2: src = d.icp_r2[level]; n = nl * 4; break;
        case RT_DBG_ICP_VAR: src = d.icp_var + level * 4; n = 12; break;
        case RT_DBG_SYSTEM: src = d.sys; n = 54 * 8; break;
        case RT_DBG_KF_WEIGHT: src = d.kf_w; n = npx * 4; break;
        case RT_DBG_MAD_HIST: src = d.madhist + 256; n = (size_t)t->prm.mad_bins * 4; break;
        default: return set_error(RT_ERR_ARG, "unknown buffer");
    }
    if (bytes < n) return set_error(RT_ERR_ARG, "buffer too small");
    HIP_TRY(hipMemcpyAsync(host, src, n, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}

extern "C" int rt_debug_set_trel(rt_tracker* t, const float* T) {
    if (!t || !T) return set_error(RT_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(t->device));
    HIP_TRY(hipMemcpyAsync(t->d.trel, T, 64, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}

extern "C" int rt_debug_prepare(rt_tracker* t) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    HIP_TRY(hipSetDevice(t->device));
    if (!ready(t)) return set_error(RT_ERR_STATE, "keyframe and current frame required");
    int rc = preprocess(t, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}

extern "C" int rt_debug_linearize(rt_tracker* t, int level, int it) {
    if (!t) return set_error(RT_ERR_ARG, "null tracker");
    if (level < 0 || level >= t->L) return set_error(RT_ERR_ARG, "bad level");
    HIP_TRY(hipSetDevice(t->device));
    SolveCfg c = solve_cfg(t);
    int m = t->mode;
    if (m == RT_PHOTOMETRIC) {
        launch_photo_iteration(t->d, level, c, FINISH_COMBINE, t->stream);
    } else if (m == RT_GEOMETRIC) {
        if (it == 0) launch_icp_select(t->d, level, t->stream);
        launch_icp_system(t->d, level, c, FINISH_COMBINE, t->stream);
    } else {
        launch_biobjective_iteration(t->d, level, c, FINISH_COMBINE, it == 0, t->stream);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(t->stream));
    return RT_OK;
}
