// C entry points over the reference's own host pose maths and trajectory
// writer, compiled from the reference sources in place (tracker/tracker/expm.c,
// basic_math.cpp, tracker/calib/ResultTxt.cpp) into
// oracle/_ref/libref.so. TEST INFRASTRUCTURE ONLY: used to write and check the
// golden vectors that pin the oracle's expm / pose functions.
#include <calib/ResultTxt.h>
#include <tracker/basic_math.h>
#include <tracker/expm.h>

extern "C" {
// expm(x, T, scale): T <- T * expm(A(x)) with the translation scaled (expm.c:2084-2090)
void ref_expm(float* x, float* T, float scale) { expm(x, T, scale); }
void ref_invert_rt4_d(double* M, double* Mi) { invertRT4(M, Mi); }
void ref_invert_rt4_f(float* M, float* Mi) { invertRT4(M, Mi); }
void ref_mat_mult4_d(double* A, double* B, double* C) { matrixMult4x4(A, B, C); }
void ref_rot2quat_f(float* m, float* q) { rot2Quaternion(m, 4, q); }
void ref_normalize_quat_f(float* q) { normalizeQuaternion(q); }
void ref_pose_distance(float* dT, double* dist, double* angle) { poseDistance(dT, dist, angle); }
int ref_inverse3x3_f(float* m, float* mi) { return inverse3x3(m, mi); }
// ResultTxt::save of n poses (calib/ResultTxt.cpp:52-76)
void ref_result_txt_save(const char* fn, int n, float* poses, int transpose, int inverse) {
    ResultTxt* r = new ResultTxt();
    r->init(fn, false);
    for (int i = 0; i < n; i++) r->addPose(poses + 16 * i);
    r->save(transpose, inverse);
    delete r;
}
}
