// Host-side launchers of the tracker kernels (rt_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rt {

// Per-level point records, structure of arrays over the N points of a level:
// {J0,J1,J2,J3} (float4), {J4,J5} (float2), reference gray (float); 28 bytes a
// point. The residual pass reads only the gray, the system pass only J.
struct JacView {
    float4* a;
    float2* b;
    float* iref;
};
__host__ __device__ inline JacView jac_view(float4* base, int N) {
    JacView v;
    v.a = base;
    v.b = reinterpret_cast<float2*>(base + N);
    v.iref = reinterpret_cast<float*>(v.b + N);
    return v;
}
inline size_t jac_bytes(int N) { return (size_t)N * 28; }

// Device buffers of one image pyramid.
struct Pyr {
    float* gray[4];
    float* z;        // level-0 depth (z < 0 valid)
};

// Device state the kernels of one tracker share.
struct Dev {
    int W, H, L, N;             // level-0 size, levels, photometric points
    const float* calib;         // 96 floats
    Pyr ref, cur;
    float* kf_w;                // fusion weights W*H
    // selection
    float* gmag;                // W*H
    unsigned* ghist;            // 256
    int* seg;                   // 512 cntA | 512 cntE | 512 startA | 512 startE | 512 mcTrim | 4 (thr, excess, totalA, total)
    int* sel;                   // N
    float4* pts;                // N
    float4* jac[4];             // per level: jac_view layout, 28 bytes a point
    // photometric iteration
    float2* res;                // N: r, zWeight
    float* wts;                 // N: robust weight of the last iteration
    unsigned* madhist;          // 512: accumulating histogram | copy of the last one
    double* part;               // photometric system partials (27 x PART_STRIDE)
    double* sys;                // 54: photometric 27 | geometric 27
    float* trel;                // 16
    // geometric
    float4* rpos[4];            // ref cloud (x,y,z,mask)
    float4* rnrm[4];            // ref normals
    float4* cpos[4];            // current cloud (x,y,z,mask)
    float* rfull;               // W*H: residual of the full-image pass
    int* icp_cnt;               // per-workgroup valid counts (W*H/1024)
    uint64_t* icp_bits;         // valid-pixel ballots, 16 words per workgroup
    int* icp_sel[4];            // per level W_l*H_l
    float* icp_r2[4];           // per level
    float* icp_var;             // 4 levels x 4 (3 rounds + spare)
    int* icp_n;                 // 4
    float* bsum;                // Student-t block sums, 3 rounds x ceil(W*H/1024)
    double* part2;              // geometric system partials (27 x PART_STRIDE)
    double* part3;              // raw geometric partials (27 x RAW_STRIDE)
    double* pose;               // device-side pose state for sequences: Tprev(16 float as double) | global(16)
    // inspection images (updateErrorImage, getDebugImage1C, getRefImage1C)
    float* dbg_err;             // W*H photometric error image
    float* dbg_ref;             // W*H normalised keyframe depth
    float* dbg_icp[4];          // per level geometric error images
    float* dbg_icp_ref[4];      // per level geometric reference images (never written, as in the reference)
};

struct SolveCfg {
    int solver;       // 0 Cholesky, 1 CG
    int mode;         // tracking mode
    float lambda;
    float scale_out;  // optScaleOut
    int robust;
    int mad_bins;
    float scale_in;   // optScaleIn
};

void launch_convert_pyramid(const Dev& d, const Pyr& p, const uint8_t* g8, const uint16_t* d16, const float* rgbd,
                            bool zero_madhist, bool zero_ghist, hipStream_t s);
void launch_reset_pose(float* T, double* pose, hipStream_t s);
void launch_fill(float* p, float v, int n, hipStream_t s);
void launch_prepare_photo_reference(const Dev& d, hipStream_t s);
void launch_ref_cloud(const Dev& d, hipStream_t s);
void launch_ref_oriented(const Dev& d, const float* xyz, const float* nrm, hipStream_t s);
void launch_ref_cloud_pyramid(const Dev& d, hipStream_t s);
void launch_cur_cloud(const Dev& d, hipStream_t s);
// How an iteration's system pass ends: combined into sys only (inspection), or
// combined, solved and the pose updated.
enum { FINISH_COMBINE = 1, FINISH_SOLVE = 2 };
// System passes: at most SYS_GROUPS workgroups of SYS_THREADS threads, each
// leaving 27 FP64 partial sums at part[k * PART_STRIDE + group].
constexpr int SYS_THREADS = 1024;
constexpr int SYS_GROUPS = 64;
constexpr int PART_STRIDE = SYS_GROUPS;
// Geometric system beside the photometric residuals: up to RAW_STRIDE
// workgroups, partials at part3[k * RAW_STRIDE + group].
constexpr int RAW_STRIDE = 256;
void launch_photo_iteration(const Dev& d, int layer, const SolveCfg& c, int finish, hipStream_t s);
void launch_icp_select(const Dev& d, int layer, hipStream_t s);
// finish is FINISH_COMBINE or FINISH_SOLVE.
void launch_icp_system(const Dev& d, int layer, const SolveCfg& c, int finish, hipStream_t s);
// One bi-objective iteration: photometric residuals, the ICP selection of the
// level (select), both systems in one launch, the combine (and solve).
void launch_biobjective_iteration(const Dev& d, int layer, const SolveCfg& c, int finish, bool select, hipStream_t s);
void launch_fuse(const Dev& d, hipStream_t s);
void launch_error_image_photo(const Dev& d, hipStream_t s);
void launch_error_image_icp(const Dev& d, hipStream_t s);
void launch_apply_increment(const Dev& d, double* pose_out, hipStream_t s);
int photo_groups(int N);

}  // namespace rt
