// Files and keyframe search around the tracker: the trajectory file, the
// tracking schedule, and the pose-distance search of similar keyframes (host
// code).
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "rt_host.hpp"
#include "rt_trajectory.h"

namespace rt {

// matrixMult4x4 for floats (basic_math.cpp:85-112): row by column, through a
// temporary, so the output may alias an input
static void mat4_mul_f(const float* A, const float* B, float* C) {
    float t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] + A[i * 4 + 2] * B[2 * 4 + j] +
                           A[i * 4 + 3] * B[3 * 4 + j];
    memcpy(C, t, sizeof(t));
}
static void transpose4(const float* M, float* R) {
    float t[16];
    for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) t[c * 4 + r] = M[r * 4 + c];
    memcpy(R, t, sizeof(t));
}

// rot2euler (basic_math.cpp:1059-1068) of a row-major 4x4 (FP64)
static void rot2euler(const double* m, double* th) {
    const double r0 = m[0], r1 = m[1], r2 = m[2], r3 = m[4], r4 = m[5], r5 = m[6], r6 = m[8], r7 = m[9], r8 = m[10];
    th[0] = atan2(r5, r8);
    const double c2 = sqrt(r0 * r0 + r1 * r1);
    th[1] = atan2(-r2, c2);
    const double s1 = sin(th[0]), c1 = cos(th[0]);
    th[2] = atan2(s1 * r6 - c1 * r3, c1 * r4 - s1 * r7);
}

// poseDistance (basic_math.cpp:1070-1094); its rad2deg (basic_math.cpp:35-37)
// takes and returns float
static void pose_distance(const float* dTf, double* dist, double* angle) {
    double dT[16];
    for (int i = 0; i < 16; i++) dT[i] = (double)dTf[i];
    const double dx = dT[3], dy = dT[7], dz = dT[11];
    *dist = (float)sqrt(dx * dx + dy * dy + dz * dz + 1e-12f);
    double th[3];
    rot2euler(dT, th);
    const double a = fabs(th[0]), b = fabs(th[1]), c = fabs(th[2]);
    const double r = (a > b) ? a : b;
    const float x = (float)((c > r) ? c : r);
    *angle = (float)(180.0 * x / M_PI);
}

// projectionDistance (KeyFrameModel.cpp:1423-1451): mean squared image
// displacement of the test object under relT, projected with K (float)
static float projection_distance(const float* relT, const float* obj, int n, const float* K) {
    float err = 0.0f;
    for (int i = 0; i < n; i++) {
        float p1[3] = {obj[3 * i], obj[3 * i + 1], obj[3 * i + 2]}, p2[3];
        p2[0] = relT[0] * p1[0] + relT[1] * p1[1] + relT[2] * p1[2] + relT[3];
        p2[1] = relT[4] * p1[0] + relT[5] * p1[1] + relT[6] * p1[2] + relT[7];
        p2[2] = relT[8] * p1[0] + relT[9] * p1[1] + relT[10] * p1[2] + relT[11];
        p1[0] /= p1[2]; p1[1] /= p1[2]; p1[2] = 1.0f;
        p2[0] /= p2[2]; p2[1] /= p2[2]; p2[2] = 1.0f;
        const float r1x = K[0] * p1[0] + K[1] * p1[1] + K[2] * p1[2], r1y = K[3] * p1[0] + K[4] * p1[1] + K[5] * p1[2];
        const float r2x = K[0] * p2[0] + K[1] * p2[1] + K[2] * p2[2], r2y = K[3] * p2[0] + K[4] * p2[1] + K[5] * p2[2];
        const float dX = r1x - r2x, dY = r1y - r2y;
        err += dX * dX + dY * dY;
    }
    return err / float(n);
}

}  // namespace rt

using namespace rt;

extern "C" int rt_trajectory_write(const char* path, int n, const float* poses, int transpose, int inverse) {
    if (!path || !poses || n < 1) return set_error(RT_ERR_ARG, "rt_trajectory_write: bad argument");
    FILE* f = fopen(path, "wb");
    if (!f) return set_error(RT_ERR_ARG, "rt_trajectory_write: cannot open the file");
    float mi[16];
    invert_rt4f(poses, mi);  // canonize (ResultTxt.cpp:52-58): poses relative to the first
    for (int i = 0; i < n; i++) {
        float m[16], mt[16], mat[16];
        mat4_mul_f(mi, poses + 16 * (size_t)i, m);
        if (inverse) invert_rt4f(m, mt); else memcpy(mt, m, sizeof(mt));
        if (transpose) transpose4(mt, mat); else memcpy(mat, mt, sizeof(mat));
        fprintf(f, "%d: %e %e %e %e %e %e %e %e %e %e %e %e %e %e %e %e\n", i + 1, mat[0], mat[4], mat[8], mat[12],
                mat[1], mat[5], mat[9], mat[13], mat[2], mat[6], mat[10], mat[14], mat[3], mat[7], mat[11], mat[15]);
    }
    return fclose(f) == 0 ? RT_OK : set_error(RT_ERR_ARG, "rt_trajectory_write: write failed");
}

extern "C" int rt_trajectory_read(const char* path, int transpose, int inverse, int max, float* poses, int* n) {
    if (!path || !n || (max > 0 && !poses)) return set_error(RT_ERR_ARG, "rt_trajectory_read: bad argument");
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(RT_ERR_ARG, "rt_trajectory_read: cannot open the file");
    int count = 0;
    char line[1024];
    while (fgets(line, sizeof(line), f)) {
        int k;
        float v[16];
        if (sscanf(line, "%d: %e %e %e %e %e %e %e %e %e %e %e %e %e %e %e %e", &k, &v[0], &v[1], &v[2], &v[3],
                   &v[4], &v[5], &v[6], &v[7], &v[8], &v[9], &v[10], &v[11], &v[12], &v[13], &v[14], &v[15]) != 17)
            continue;
        if (count >= max) { count++; continue; }
        float mat[16], mt[16];
        transpose4(v, mat);  // the line holds the matrix column by column
        if (transpose) transpose4(mat, mt); else memcpy(mt, mat, sizeof(mt));
        if (inverse) invert_rt4f(mt, poses + 16 * (size_t)count); else memcpy(poses + 16 * (size_t)count, mt, sizeof(mt));
        count++;
    }
    fclose(f);
    *n = count;
    return count > max ? set_error(RT_ERR_ARG, "rt_trajectory_read: more poses than max") : RT_OK;
}

extern "C" int rt_schedule_read(const char* path, int max, rt_schedule_entry* out, int* n) {
    if (!path || !n || (max > 0 && !out)) return set_error(RT_ERR_ARG, "rt_schedule_read: bad argument");
    FILE* f = fopen(path, "rb");
    if (!f) return set_error(RT_ERR_ARG, "rt_schedule_read: cannot open the file");
    int count = 0;
    char line[1024];
    while (fgets(line, sizeof(line), f)) {
        int frame = 0;
        char name[512] = {0};
        float err = 0.0f;
        if (sscanf(line, "%d %511s %e", &frame, name, &err) < 2) continue;
        if (count >= max) { count++; continue; }
        rt_schedule_entry& e = out[count++];
        e.frame = frame; e.error = err;
        if (strcmp(name, "geometric") == 0) { e.mode = RT_GEOMETRIC; e.lambda = 1.0f; }
        else if (strcmp(name, "photometric") == 0) { e.mode = RT_PHOTOMETRIC; e.lambda = 0.0f; }
        else if (strncmp(name, "biobjective", strlen("biobjective")) == 0) {
            e.mode = RT_BIOBJECTIVE;
            // the reference reads the number after "biobjective-lambda"
            e.lambda = strlen(name) >= strlen("biobjective-lambda") ? (float)atof(name + strlen("biobjective-lambda")) : 0.0f;
        } else { e.mode = -1; e.lambda = 0.0f; }
    }
    fclose(f);
    *n = count;
    return count > max ? set_error(RT_ERR_ARG, "rt_schedule_read: more entries than max") : RT_OK;
}

extern "C" int rt_schedule_lookup(const rt_schedule_entry* entries, int n, int frame, int* mode, float* lambda) {
    if (!mode || !lambda || n < 0 || (n > 0 && !entries)) return set_error(RT_ERR_ARG, "rt_schedule_lookup: bad argument");
    if (n == 0) { *mode = RT_GEOMETRIC; *lambda = 1.0f; return RT_OK; }
    int best = 0, bd = abs(entries[0].frame - frame);
    for (int i = 1; i < n; i++) {
        const int d = abs(entries[i].frame - frame);
        if (d < bd) { bd = d; best = i; }
    }
    *mode = entries[best].mode;
    *lambda = entries[best].lambda;
    return RT_OK;
}

extern "C" int rt_pose_distance(const float* dT, double* dist, double* angle_deg) {
    if (!dT || !dist || !angle_deg) return set_error(RT_ERR_ARG, "rt_pose_distance: bad argument");
    pose_distance(dT, dist, angle_deg);
    return RT_OK;
}

extern "C" int rt_grid_object(int nx, int ny, int nz, float size, float view_min, float z_range, float fov_x_deg,
                              float* out) {
    if (!out || nx < 1 || ny < 1 || nz < 1) return set_error(RT_ERR_ARG, "rt_grid_object: bad argument");
    const float sx = (2 * size) / nx, sy = (2 * size) / ny;
    float sz = 0.0f;
    if (nz == 1) view_min += z_range / 2.0f;
    else sz = z_range / float(nz - 1);
    const float tf = tanf(fov_x_deg * 3.141592653f / 180.0f);
    size_t p = 0;
    for (int k = 0; k < nz; k++) {
        float scale = 1.0f;
        if (k > 0) {
            const float z0 = -view_min, zk = -view_min - sz * k;
            scale = (zk * tf) / (z0 * tf);
        }
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++, p++) {
                out[3 * p + 0] = (-size + i * sx) * scale;
                out[3 * p + 1] = (-size + j * sy) * scale;
                out[3 * p + 2] = -view_min - sz * k;
            }
    }
    return RT_OK;
}

extern "C" int rt_keyframes_similar(const float* T, int nkeys, const float* key_poses, float max_dist, float max_angle,
                                    const float* K, const float* grid, int nvertices, int* order, float* weights,
                                    int* count) {
    if (!T || !count || nkeys < 0 || (nkeys > 0 && (!key_poses || !order || !weights)) || !K || !grid || nvertices < 1)
        return set_error(RT_ERR_ARG, "rt_keyframes_similar: bad argument");
    float iT[16];
    invert_rt4f(T, iT);
    std::vector<int> close;
    std::vector<float> w;
    float wmin = FLT_MAX, wmax = FLT_MIN;  // the reference's initial values (KeyFrameModel.cpp:1580)
    for (int k = 0; k < nkeys; k++) {
        float rel[16];
        mat4_mul_f(iT, key_poses + 16 * (size_t)k, rel);  // getRelativeTransform (TrackFrame.cpp:542-549)
        double d, a;
        pose_distance(rel, &d, &a);
        if (!((float)d < max_dist && (float)a < max_angle)) continue;
        const float v = projection_distance(rel, grid, nvertices, K);
        if (v < wmin) wmin = v;
        if (v > wmax) wmax = v;
        close.push_back(k);
        w.push_back(v);
    }
    // bubbleSort (KeyFrameModel.cpp:1530-1544): ascending, equal ones keep their order
    std::vector<int> idx(close.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] < w[b]; });
    for (size_t i = 0; i < idx.size(); i++) {
        order[i] = close[idx[i]];
        weights[i] = 1.0f - (w[idx[i]] - wmin) / (wmax - wmin);
    }
    *count = (int)close.size();
    return RT_OK;
}
