// Host-side helpers shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prrtc.h"
#include "device_common.h"

namespace prrtc {

void set_error(const char *fmt, ...);

struct HipError {
    hipError_t code;
    const char *where;
};

#define PRRTC_HIP(call)                                                   \
    do {                                                                  \
        hipError_t _e = (call);                                           \
        if (_e != hipSuccess) throw ::prrtc::HipError{_e, #call};         \
    } while (0)

// Run `body`, turning exceptions into status codes and the thread's message.
template <class F>
int guarded(const char *fn, F &&body) {
    try {
        return body();
    } catch (const HipError &e) {
        set_error("%s: %s failed: %s", fn, e.where, hipGetErrorString(e.code));
        return PRRTC_EDEVICE;
    } catch (const std::bad_alloc &) {
        set_error("%s: host allocation failed", fn);
        return PRRTC_ENOMEM;
    }
}

inline int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// Device buffer owned by the host code.
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) PRRTC_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

int robot_dof(int robot);
int robot_num_spheres(int robot);
bool scene_valid(const prrtc_scene *s);
size_t scene_packed_floats(const prrtc_scene *s);
void scene_pack(const prrtc_scene *s, float *out);
int num_cus();

}  // namespace prrtc
