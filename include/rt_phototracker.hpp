// rt_phototracker.hpp -- the reference's PhotoTracker host interface over the C ABI.
//
// Header-only C++ class with the method names, argument meaning and error
// behaviour of PhotoTracker (tracker/tracker/phototracker.h:33-82), so that a
// caller of the reference API keeps its call sites. Differences, all forced by
// the boundary:
//  - frames arrive as 8-bit gray + 16-bit depth (mm) buffers, or as the
//    reference's float4 RGBD image, instead of a cudaArray;
//  - there is no tracker thread and no barrier (phototracker.cpp:691-728): the
//    caller's thread runs one frame with track() (readyToTrack, gpuPreProcess,
//    gpuEstimateIncrement and applyIncrement in one call);
//  - setReferencePoints takes the keyframe frame and, for the filtered
//    reference, the ray-cast point and normal maps as device float4 images
//    (the TSDF model of rt_tsdf.h makes them).
// Returned pose pointers are borrowed arrays owned by the object, valid until
// the next call that changes the pose, as in the reference. Failures print
// rt_last_error() to stderr and return false / leave state unchanged; the
// reference prints and returns on bad arguments (phototracker.cpp:256-266).
#pragma once
#include <stdio.h>

#include "rt_tracker.h"

namespace rt {

enum TrackMode { GEOMETRIC = RT_GEOMETRIC, PHOTOMETRIC = RT_PHOTOMETRIC, BIOBJECTIVE = RT_BIOBJECTIVE };

class PhotoTracker {
public:
    // initializePhotoTracker(deviceID, texWidth, texHeight, calib) (phototracker.cpp:900-913)
    static PhotoTracker* create(int deviceID, int width, int height, const float calib[RT_CALIB_SIZE],
                                const rt_params* params = nullptr) {
        rt_params p;
        if (params) p = *params;
        else rt_default_params(&p);
        rt_tracker* h = nullptr;
        if (rt_create(deviceID, width, height, calib, &p, &h) != RT_OK) {
            report("initializePhotoTracker");
            return nullptr;
        }
        return new PhotoTracker(h, deviceID);
    }
    ~PhotoTracker() { rt_destroy(h_); }
    PhotoTracker(const PhotoTracker&) = delete;
    PhotoTracker& operator=(const PhotoTracker&) = delete;

    int getDeviceID() const { return device_; }
    rt_tracker* handle() { return h_; }

    double* getGlobalPose() {
        if (rt_get_global_pose(h_, global_) != RT_OK) report("getGlobalPose");
        return global_;
    }
    float* getRelativePose() {
        if (rt_get_relative_pose(h_, relative_) != RT_OK) report("getRelativePose");
        return relative_;
    }
    double* getIncrement() {
        if (rt_get_increment(h_, increment_) != RT_OK) report("getIncrement");
        return increment_;
    }
    void resetTracking(TrackMode mode = GEOMETRIC) { check(rt_reset_tracking(h_, mode), "resetTracking"); }
    void clearPose() { check(rt_clear_pose(h_), "clearPose"); }
    void resetToTransform(double* T) { check(rt_reset_to_transform(h_, T), "resetToTransform"); }

    void setCurrentImage(const uint8_t* gray, const uint16_t* depth, bool onDevice) {
        check(rt_set_current(h_, gray, depth, onDevice ? 1 : 0), "setCurrentImage");
    }
    void setCurrentImage(const float* rgbd, bool onDevice) {
        check(rt_set_current_rgbd(h_, rgbd, onDevice ? 1 : 0), "setCurrentImage");
    }
    // true when the keyframe was replaced (referenceOutdated or forced)
    bool setReferencePoints(const uint8_t* gray, const uint16_t* depth, bool onDevice, TrackMode suggestedMethod,
                            float lambda = 0.0f, bool force = false) {
        int rc = rt_set_reference(h_, gray, depth, onDevice ? 1 : 0, suggestedMethod, lambda, force ? 1 : 0);
        if (rc < 0) report("setReferencePoints");
        return rc == 1;
    }
    bool setReferencePoints(const float* rgbd, bool onDevice, TrackMode suggestedMethod, float lambda = 0.0f,
                            bool force = false) {
        int rc = rt_set_reference_rgbd(h_, rgbd, onDevice ? 1 : 0, suggestedMethod, lambda, force ? 1 : 0);
        if (rc < 0) report("setReferencePoints");
        return rc == 1;
    }
    // with the ray-cast maps of a filtered reference (device float4 images)
    bool setReferencePoints(const float* xyz, const float* normals, const uint8_t* gray, const uint16_t* depth,
                            bool onDevice, TrackMode suggestedMethod, float lambda = 0.0f, bool force = false) {
        int rc = rt_set_reference_points(h_, xyz, normals, gray, depth, onDevice ? 1 : 0, suggestedMethod, lambda,
                                         force ? 1 : 0);
        if (rc < 0) report("setReferencePoints");
        return rc == 1;
    }
    bool setReferencePoints(const float* xyz, const float* normals, const float* rgbd, bool onDevice,
                            TrackMode suggestedMethod, float lambda = 0.0f, bool force = false) {
        int rc = rt_set_reference_points_rgbd(h_, xyz, normals, rgbd, onDevice ? 1 : 0, suggestedMethod, lambda,
                                              force ? 1 : 0);
        if (rc < 0) report("setReferencePoints");
        return rc == 1;
    }
    // updateErrorImage / getDebugImage1C / getRefImage1C, copied to host memory
    void updateErrorImage() { check(rt_update_error_image(h_), "updateErrorImage"); }
    bool getDebugImage1C(int layer, float* out, size_t count) {
        int rc = rt_get_debug_image(h_, layer, out, count);
        if (rc < 0) report("getDebugImage1C");
        return rc == RT_OK;
    }
    bool getRefImage1C(int layer, float* out, size_t count) {
        int rc = rt_get_ref_image(h_, layer, out, count);
        if (rc < 0) report("getRefImage1C");
        return rc == RT_OK;
    }
    void setFilteredReference(bool flag) { check(rt_set_filtered_reference(h_, flag ? 1 : 0), "setFilteredReference"); }
    bool getFilteredReference() const { return rt_get_filtered_reference(h_) == 1; }
    // One frame of the tracker thread's loop body (phototracker.cpp:704-718);
    // false when not ready to track (no keyframe or no current frame).
    bool track() {
        int rc = rt_track(h_);
        if (rc < 0) report("track");
        return rc == 1;
    }

    void setIterationCounts(const int* nIter) { check(rt_set_iteration_counts(h_, nIter), "setIterationCounts"); }
    void getIterationCounts(int* nIter) { check(rt_get_iteration_counts(h_, nIter), "getIterationCounts"); }
    void setMode(TrackMode mode) { check(rt_set_mode(h_, mode), "setMode"); }
    TrackMode getMode() { return (TrackMode)rt_get_mode(h_); }
    void setLambda(float lambda) { check(rt_set_lambda(h_, lambda), "setLambda"); }
    float getLambda() { return rt_get_lambda(h_); }
    void setReferenceUpdating(bool flag) { check(rt_set_reference_updating(h_, flag ? 1 : 0), "setReferenceUpdating"); }
    void setIncrementalMode(bool flag) { check(rt_set_incremental_mode(h_, flag ? 1 : 0), "setIncrementalMode"); }
    void signalCalibrationUpdate(float* calibDataExt) {
        check(rt_signal_calibration_update(h_, calibDataExt), "signalCalibrationUpdate");
    }
    int numReferenceUpdates() { return rt_num_reference_updates(h_); }
    int getLayers() { return rt_get_layers(h_); }
    // listSelected: current and reference image coordinates (x,y pairs) and
    // weights of the selected points; capacity entries per output.
    void listSelected(float* selected, float* refSelected, float* weight, int capacity, int* nSelect) {
        check(rt_list_selected(h_, selected, refSelected, weight, capacity, nSelect), "listSelected");
    }
    // Keyframe fusion of the current frame, and re-preparation of the fused keyframe.
    void fuseCurrent() { check(rt_fuse_current(h_), "fuseCurrent"); }
    void refreshKeyframe() { check(rt_refresh_keyframe(h_), "refreshKeyframe"); }
    void sync() { check(rt_synchronize(h_), "sync"); }

private:
    PhotoTracker(rt_tracker* h, int device) : h_(h), device_(device) {}
    static void report(const char* what) {
        const char* e = rt_last_error();
        fprintf(stderr, "PhotoTracker::%s: %s\n", what, e ? e : "error");
    }
    static void check(int rc, const char* what) {
        if (rc < 0) report(what);
    }
    rt_tracker* h_;
    int device_;
    double global_[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double increment_[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    float relative_[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
};

}  // namespace rt
