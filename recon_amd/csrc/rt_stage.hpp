// Page-locked host staging for asynchronous uploads.
//
// An upload queued with hipMemcpyAsync may read its host source after the
// call that queued it has returned. Tables built on the host for one launch
// (pair descriptors, frame tables, flags, initial poses) therefore go through
// a persistent pinned buffer: get() waits until the copy that last read the
// buffer has completed (an event recorded after it), so the buffer is never
// rewritten while a queued copy may still read it, and the host data it held
// outlives the queuing call.
#pragma once
#include <hip/hip_runtime.h>

#include <string.h>

namespace rt {

struct HostStage {
    void* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;

    // host memory of at least `bytes`, free of queued readers; nullptr on failure
    void* get(size_t bytes) {
        if (pending) {
            if (hipEventSynchronize(ev) != hipSuccess) return nullptr;
            pending = false;
        }
        if (bytes > cap) {
            if (p) hipHostFree(p);
            p = nullptr;
            cap = 0;
            if (hipHostMalloc(&p, bytes, 0) != hipSuccess) { p = nullptr; return nullptr; }
            cap = bytes;
        }
        if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) { ev = nullptr; return nullptr; }
        return p;
    }
    // after the uploads from the buffer have been queued on `s`
    bool queued(hipStream_t s) {
        if (hipEventRecord(ev, s) != hipSuccess) return false;
        pending = true;
        return true;
    }
    // copy `bytes` of host data through the stage to `dst` on `s`
    bool upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
        void* h = get(bytes);
        if (!h) return false;
        memcpy(h, src, bytes);
        return send(dst, bytes, s);
    }
    // copy the first `bytes` of the stage, written in place after get(), to `dst` on `s`
    bool send(void* dst, size_t bytes, hipStream_t s) {
        return hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, s) == hipSuccess && queued(s);
    }
    void release() {
        if (pending) hipEventSynchronize(ev);
        if (p) hipHostFree(p);
        if (ev) hipEventDestroy(ev);
        p = nullptr; ev = nullptr; cap = 0; pending = false;
    }
};

}  // namespace rt
