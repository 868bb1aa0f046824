// Host-side pose maths of the tracker (CPU, FP64 where the reference uses it).
#pragma once
#include <math.h>
#include <string.h>

namespace rt {

int set_error(int code, const char* msg);

// invertRT4 (tracker/tracker/basic_math.cpp:335-367)
template <class M>
inline void invert_rt4(const M* A, M* Ai) {
    M t[16];
    t[0] = A[0]; t[1] = A[4]; t[2] = A[8];   t[3] = -(A[0] * A[3] + A[4] * A[7] + A[8] * A[11]);
    t[4] = A[1]; t[5] = A[5]; t[6] = A[9];   t[7] = -(A[1] * A[3] + A[5] * A[7] + A[9] * A[11]);
    t[8] = A[2]; t[9] = A[6]; t[10] = A[10]; t[11] = -(A[2] * A[3] + A[6] * A[7] + A[10] * A[11]);
    t[12] = 0; t[13] = 0; t[14] = 0; t[15] = 1;
    memcpy(Ai, t, sizeof(t));
}
inline void invert_rt4f(const float* A, float* Ai) { invert_rt4<float>(A, Ai); }

// matrixMult4x4 (basic_math.cpp:183-208)
inline void mat4_mul(const double* A, const double* B, double* C) {
    double t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            t[i * 4 + j] = A[i * 4 + 0] * B[0 * 4 + j] + A[i * 4 + 1] * B[1 * 4 + j] + A[i * 4 + 2] * B[2 * 4 + j] +
                           A[i * 4 + 3] * B[3 * 4 + j];
    memcpy(C, t, sizeof(t));
}

// PhotoTracker::applyIncrement (phototracker.cpp:514-531): trel is TrelDev
// (reference points -> current view); T is the host relative pose (its
// inverse), Tinc the increment since the previous frame, global the
// accumulated pose.
inline void apply_increment(const float* trel, float* T, double* Tinc, double* global) {
    double Tprev[16], iTprev[16], iT[16], Td[16];
    for (int i = 0; i < 16; i++) Tprev[i] = T[i];
    invert_rt4(Tprev, iTprev);
    for (int i = 0; i < 16; i++) iT[i] = trel[i];
    invert_rt4(iT, Td);
    for (int i = 0; i < 16; i++) T[i] = (float)Td[i];
    mat4_mul(iTprev, Td, Tinc);
    mat4_mul(global, Tinc, global);
}

// rot2Quaternion (basic_math.cpp:943-980, row stride 4)
inline void rot2quat(const float* m, float* q) {
    float tr = m[0] + m[5] + m[10];
    if (tr > 0) {
        float s = 0.5f / sqrtf(tr + 1.0f);
        q[0] = 0.25f / s; q[1] = (m[9] - m[6]) * s; q[2] = (m[2] - m[8]) * s; q[3] = (m[4] - m[1]) * s;
    } else if (m[0] > m[5] && m[0] > m[10]) {
        float s = 2.0f * sqrtf(1.0f + m[0] - m[5] - m[10]);
        q[0] = (m[9] - m[6]) / s; q[1] = 0.25f * s; q[2] = (m[1] + m[4]) / s; q[3] = (m[2] + m[8]) / s;
    } else if (m[5] > m[10]) {
        float s = 2.0f * sqrtf(1.0f + m[5] - m[0] - m[10]);
        q[0] = (m[2] - m[8]) / s; q[1] = (m[1] + m[4]) / s; q[2] = 0.25f * s; q[3] = (m[6] + m[9]) / s;
    } else {
        float s = 2.0f * sqrtf(1.0f + m[10] - m[0] - m[5]);
        q[0] = (m[4] - m[1]) / s; q[1] = (m[2] + m[8]) / s; q[2] = (m[6] + m[9]) / s; q[3] = 0.25f * s;
    }
}

// rad2deg (basic_math.cpp:35-37; M_PI is the double constant of math.h there)
inline float rad2deg_f(float x) { return (float)(180.0 * (double)x / 3.14159265358979323846); }

// poseDifferenceExceeded of PhotoTracker (phototracker.cpp:268-290) and of
// ShapeTracker (shapetracker.cpp:168-186, identity test at 1e-8).
inline bool pose_difference_exceeded(const float* dT, float tthr, float athr, bool shape) {
    double dx = dT[3], dy = dT[7], dz = dT[11];
    float dist = (float)sqrt(dx * dx + dy * dy + dz * dz + 1e-12f);
    if (dist > tthr) return true;
    float angle = 0;
    double eps = shape ? 1e-8 : 1e-5;
    if ((fabs(dT[0] - 1.0f) > eps) || (fabs(dT[5] - 1.0f) > eps) || (fabs(dT[10] - 1.0f) > eps)) {
        float q[4];
        rot2quat(dT, q);
        float len = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3] + 1e-7f;
        len = (float)sqrt((double)len);
        for (int i = 0; i < 4; i++) q[i] /= len;
        double ca = (double)q[0];
        if (shape) angle = rad2deg_f((float)acos(ca)) * 2.0f;
        else angle = (float)(acos(ca) * 2.0f * 180.0f / 3.141592653f);
    }
    return angle > athr;
}

// A pinhole calibration block in the reference convention
// (Calibration::setupCalibDataBuffer, tracker/calib/calib.cpp:74-93):
// K = [[-fx,0,cx],[0,fy,cy],[0,0,1]] for both cameras, its inverse, no lens
// distortion, TLR = TRL = identity, depth window [zmin, zmax].
inline int calib_pinhole(float fx, float fy, float cx, float cy, float zmin, float zmax, float* out) {
    for (int i = 0; i < 96; i++) out[i] = 0.0f;
    float K[9] = {-fx, 0, cx, 0, fy, cy, 0, 0, 1};
    // inverse of an upper-triangular intrinsic matrix
    float iK[9] = {1.0f / K[0], 0, -K[2] / K[0], 0, 1.0f / K[4], -K[5] / K[4], 0, 0, 1};
    for (int i = 0; i < 9; i++) {
        out[0 + i] = K[i]; out[9 + i] = iK[i];
        out[43 + i] = K[i]; out[52 + i] = iK[i];
    }
    for (int i = 0; i < 16; i++) {
        float v = (i % 5 == 0) ? 1.0f : 0.0f;
        out[26 + i] = v; out[64 + i] = v;
    }
    out[62] = zmin;
    out[63] = zmax;
    return 0;
}

}  // namespace rt
