// Edge validation, pose checks and FK dumps on gfx950, plus the C ABI for
// them (include/prrtc.h).
//
// Edge validation replaces the per-edge block of the reference planner
// (src/planning/pRRTC.cu:430-445): there one CUDA thread per interpolated
// pose runs fkcc<Robot> and a shared-memory atomicOr folds the verdict.
// Here one wave64 validates one edge: each pass puts up to 64 poses on the
// lanes (poses interleaved across passes so the first pass already spans the
// whole edge), the robot's check sequence runs with the obstacle tables in
// LDS, and a ballot after every check group ends the pass -- and the edge --
// at the first collision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "host_common.h"

namespace prrtc {

// ------------------------------------------------------------------ kernels

template <class Robot>
__global__ __launch_bounds__(kWave) void k_fk_spheres(const float *__restrict__ poses, int n,
                                                      float *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float q[Robot::DOF];
#pragma unroll
    for (int d = 0; d < Robot::DOF; d++) q[d] = poses[(size_t)i * Robot::DOF + d];
    Robot::fk_spheres(q, out + (size_t)i * Robot::NSPHERES * 3);
}

template <class Robot, bool LDS>
__global__ __launch_bounds__(kWave) void k_check_poses(const float *__restrict__ scene, int scene_f4,
                                                       const float *__restrict__ poses, int n,
                                                       uint8_t *__restrict__ out) {
    extern __shared__ float4 lds_scene[];
    CheckCtx ctx;
    ctx.independent = true;
    ctx.key = 0;
    bind_scene<LDS>(scene, scene_f4, lds_scene, ctx.obs);
    for (int base = blockIdx.x * kWave; base < n; base += gridDim.x * kWave) {
        const int i = base + threadIdx.x;
        const int src = i < n ? i : n - 1;
        float q[Robot::DOF];
#pragma unroll
        for (int d = 0; d < Robot::DOF; d++) q[d] = poses[(size_t)src * Robot::DOF + d];
        ctx.coll = false;
        ctx.dead = i >= n;
        Robot::check(q, ctx);
        if (i < n) out[i] = ctx.coll ? 0 : 1;
    }
}

// One wave per edge (grid-stride over edges).  Pass p of P = ceil(g/64)
// gives lane l the pose k = l*P + p; pose k is a + (k+1) * ((b-a)/g).
template <class Robot, bool LDS>
__global__ __launch_bounds__(kWave) void k_validate_edges(const float *__restrict__ scene, int scene_f4,
                                                          const float *__restrict__ edges, int n, int g,
                                                          uint8_t *__restrict__ out) {
    extern __shared__ float4 lds_scene[];
    constexpr int D = Robot::DOF;
    CheckCtx ctx;
    ctx.independent = false;
    bind_scene<LDS>(scene, scene_f4, lds_scene, ctx.obs);
    const int lane = threadIdx.x;
    const int passes = (g + kWave - 1) / kWave;
    const float gf = (float)g;
    // Workgroups reach the XCDs round-robin (workgroup b runs on XCD b % 8),
    // and each XCD has its own L2.  Giving the workgroups of one XCD one
    // contiguous range of edges keeps every cache line of the edge list in
    // one L2 (edges are 4*2*D bytes and straddle lines).
    int lo = 0, hi = n, first = blockIdx.x, stride = gridDim.x;
    if ((int)gridDim.x >= kXcds) {
        const int x = blockIdx.x % kXcds;
        lo = (int)((long long)n * x / kXcds);
        hi = (int)((long long)n * (x + 1) / kXcds);
        first = blockIdx.x / kXcds;
        stride = ((int)gridDim.x - x + kXcds - 1) / kXcds;   // workgroups on this XCD
    }
    for (int e = lo + first; e < hi; e += stride) {
        const float *a = edges + (size_t)e * 2 * D;
        const float *b = a + D;
        float av[D], dv[D];
#pragma unroll
        for (int d = 0; d < D; d++) {
            av[d] = a[d];
            dv[d] = (b[d] - av[d]) / gf;
        }
        bool bad = false;
        for (int p = 0; p < passes && !bad; p++) {
            const int k = lane * passes + p;
            const float kk = (float)((k < g ? k : g - 1) + 1);
            float q[D];
#pragma unroll
            for (int d = 0; d < D; d++) q[d] = fmaf(kk, dv[d], av[d]);
            ctx.coll = false;
            ctx.dead = k >= g;
            ctx.key = 0;
            Robot::check(q, ctx);
            bad = ballot(ctx.coll) != 0;
        }
        if (lane == 0) out[e] = bad ? 0 : 1;
    }
}

// ------------------------------------------------------------------ helpers

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int robot_dof(int robot) {
    switch (robot) {
#define PRRTC_DOF_CASE(ID, R) case ID: return R::DOF;
        PRRTC_FOR_EACH_ROBOT(PRRTC_DOF_CASE)
#undef PRRTC_DOF_CASE
        default: return -1;
    }
}

int robot_num_spheres(int robot) {
    switch (robot) {
#define PRRTC_NSPH_CASE(ID, R) case ID: return R::NSPHERES;
        PRRTC_FOR_EACH_ROBOT(PRRTC_NSPH_CASE)
#undef PRRTC_NSPH_CASE
        default: return -1;
    }
}

bool scene_valid(const prrtc_scene *s) {
    if (!s) return false;
    const int n[5] = {s->n_spheres, s->n_capsules, s->n_z_capsules, s->n_cuboids, s->n_z_cuboids};
    const float *p[5] = {s->spheres, s->capsules, s->z_capsules, s->cuboids, s->z_cuboids};
    const int w[5] = {4, 8, 8, 15, 15};
    for (int t = 0; t < 5; t++) {
        if (n[t] < 0 || (n[t] > 0 && !p[t])) return false;
        for (int i = 0; i < n[t] * w[t]; i++)
            if (!std::isfinite(p[t][i])) return false;
    }
    return true;
}

// Cull radius of a cuboid with half-extents h: its half-diagonal |h| (the box
// lies inside the ball of that radius about its centre), padded by 2^-16
// relative and 1 mm absolute.  A robot sphere (radius r) whose centre is
// farther than r + cull from the box centre is at least 1 mm from the box,
// so the box test's value is positive by a margin far above the rounding
// of either computation: skipping the test cannot change a verdict.  A
// non-finite extent gives a non-finite radius, which never culls.
static float cuboid_cull_radius(const float *h) {
    const double R = std::sqrt((double)h[0] * h[0] + (double)h[1] * h[1] + (double)h[2] * h[2]);
    return (float)(R * (1.0 + 1.0 / 65536.0) + 1e-3);
}

size_t scene_packed_floats(const prrtc_scene *s) {
    return 8 + 4 * (size_t)scene_payload_f4(s->n_spheres, s->n_capsules, s->n_z_capsules, s->n_cuboids,
                                            s->n_z_cuboids);
}

void scene_pack(const prrtc_scene *s, float *out) {
    const int nf4 = scene_payload_f4(s->n_spheres, s->n_capsules, s->n_z_capsules, s->n_cuboids, s->n_z_cuboids);
    int32_t hdr[8] = {s->n_spheres, s->n_capsules, s->n_z_capsules, s->n_cuboids, s->n_z_cuboids, nf4, 0, 0};
    std::memcpy(out, hdr, sizeof hdr);
    float *o = out + 8;
    for (int i = 0; i < s->n_spheres; i++, o += 4) std::memcpy(o, s->spheres + 4 * i, 4 * sizeof(float));
    auto caps = [&](const float *c, int n) {
        for (int i = 0; i < n; i++, o += 8) {
            const float *r = c + 8 * i;  // x1 y1 z1 xv yv zv r rdv
            const float v[8] = {r[0], r[1], r[2], r[6], r[3], r[4], r[5], r[7]};
            std::memcpy(o, v, sizeof v);
        }
    };
    caps(s->capsules, s->n_capsules);
    caps(s->z_capsules, s->n_z_capsules);
    auto cubs = [&](const float *c, int n) {
        for (int i = 0; i < n; i++, o += 16) {
            const float *r = c + 15 * i;  // x y z a1 a2 a3 r1 r2 r3
            const float v[16] = {r[0], r[1], r[2], r[12], r[3], r[4], r[5], r[13],
                                 r[6], r[7], r[8], r[14], r[9], r[10], r[11], cuboid_cull_radius(r + 12)};
            std::memcpy(o, v, sizeof v);
        }
    };
    cubs(s->cuboids, s->n_cuboids);
    cubs(s->z_cuboids, s->n_z_cuboids);
}

int num_cus() {
    static int cus = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    });
    return cus;
}

template <class... KArgs, class... Args>
static void launch_scene_kernel(void (*k)(KArgs...), dim3 grid, size_t scene_f4, hipStream_t st, Args... args) {
    hipLaunchKernelGGL(k, grid, dim3(kWave), scene_f4 * sizeof(float4), st, args...);
    PRRTC_HIP(hipGetLastError());
}

template <class Robot>
static void launch_validate(const float *d_scene, size_t scene_f4, const float *d_edges, int n, int g,
                            uint8_t *d_out, hipStream_t st) {
    const int grid = std::max(1, std::min(n, num_cus() * 16));
    launch_scene_kernel(k_validate_edges<Robot, true>, dim3(grid), scene_f4, st, d_scene, (int)scene_f4, d_edges,
                        n, g, d_out);
}

template <class Robot>
static void launch_check(const float *d_scene, size_t scene_f4, const float *d_poses, int n, uint8_t *d_out,
                         hipStream_t st) {
    const int blocks = std::max(1, std::min((n + kWave - 1) / kWave, num_cus() * 16));
    launch_scene_kernel(k_check_poses<Robot, true>, dim3(blocks), scene_f4, st, d_scene, (int)scene_f4, d_poses,
                        n, d_out);
}

// Scenes are staged whole into LDS: at most kMaxLdsScene bytes of tables
// (4096 float4: e.g. 1024 cuboids or 4096 spheres).
static bool scene_fits(size_t scene_f4) { return scene_f4 * sizeof(float4) <= kMaxLdsScene; }

template <class Robot>
static void launch_fk(const float *d_poses, int n, float *d_out, hipStream_t st) {
    hipLaunchKernelGGL(k_fk_spheres<Robot>, dim3((n + kWave - 1) / kWave), dim3(kWave), 0, st, d_poses, n, d_out);
    PRRTC_HIP(hipGetLastError());
}

// FN<Robot>(args...) for the robot id, over every robot of
// PRRTC_FOR_EACH_ROBOT (the arguments reach each case through a lambda)
#define PRRTC_DISPATCH_CASE(ID, R) \
    case ID: call_(RobotTag<R>{}); break;
#define PRRTC_DISPATCH(robot, FN, ...)                                              \
    do {                                                                            \
        auto call_ = [&](auto tag) { FN<typename decltype(tag)::type>(__VA_ARGS__); }; \
        switch (robot) { PRRTC_FOR_EACH_ROBOT(PRRTC_DISPATCH_CASE) }                 \
    } while (0)

static bool all_finite(const float *p, size_t n) {
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(p[i])) return false;
    return true;
}

}  // namespace prrtc

using namespace prrtc;

// ------------------------------------------------------------------ C ABI

extern "C" {

int prrtc_version(void) { return 1; }

const char *prrtc_last_error(void) { return g_error.c_str(); }

void prrtc_settings_default(prrtc_settings *s) {
    if (!s) return;
    // pRRTC_settings.hh:4-16
    s->max_samples = 1000000;
    s->max_iters = 1000000;
    s->num_new_configs = 600;
    s->granularity = 64;
    s->range = 0.5f;
    s->balance = 1;
    s->tree_ratio = 0.5f;
    s->dynamic_domain = 1;
    s->dd_alpha = 0.0001f;
    s->dd_radius = 4.0f;
    s->dd_min_radius = 1.0f;
    s->seed = 1;
}

int prrtc_robot_dof(int robot) { return robot_dof(robot); }

int prrtc_robot_num_spheres(int robot) { return robot_num_spheres(robot); }

size_t prrtc_scene_packed_floats(const prrtc_scene *scene) {
    if (!scene) return 0;
    return scene_packed_floats(scene);
}

int prrtc_scene_pack(const prrtc_scene *scene, float *out, size_t cap_floats) {
    if (!scene_valid(scene) || !out) {
        set_error("prrtc_scene_pack: invalid scene or output");
        return PRRTC_EINVAL;
    }
    if (cap_floats < scene_packed_floats(scene)) {
        set_error("prrtc_scene_pack: output holds %zu floats, %zu needed", cap_floats, scene_packed_floats(scene));
        return PRRTC_EOVERFLOW;
    }
    scene_pack(scene, out);
    return PRRTC_OK;
}

int prrtc_fk_spheres(int robot, const float *poses, int n, float *out_centres) {
    return guarded("prrtc_fk_spheres", [&]() -> int {
        const int D = robot_dof(robot), S = robot_num_spheres(robot);
        if (D < 0 || n < 0 || (n > 0 && (!poses || !out_centres))) {
            set_error("prrtc_fk_spheres: invalid argument");
            return PRRTC_EINVAL;
        }
        if (n == 0) return PRRTC_OK;
        DevBuf<float> dq((size_t)n * D), dc((size_t)n * S * 3);
        PRRTC_HIP(hipMemcpy(dq.p, poses, sizeof(float) * n * D, hipMemcpyHostToDevice));
        PRRTC_DISPATCH(robot, launch_fk, dq.p, n, dc.p, (hipStream_t)0);
        PRRTC_HIP(hipMemcpy(out_centres, dc.p, sizeof(float) * (size_t)n * S * 3, hipMemcpyDeviceToHost));
        return PRRTC_OK;
    });
}

int prrtc_check_poses(int robot, const prrtc_scene *scene, const float *poses, int n, uint8_t *out_free) {
    return guarded("prrtc_check_poses", [&]() -> int {
        const int D = robot_dof(robot);
        if (D < 0 || n < 0 || !scene_valid(scene) || (n > 0 && (!poses || !out_free))) {
            set_error("prrtc_check_poses: invalid argument");
            return PRRTC_EINVAL;
        }
        if (n == 0) return PRRTC_OK;
        if (!all_finite(poses, (size_t)n * D)) {
            set_error("prrtc_check_poses: non-finite pose");
            return PRRTC_EINVAL;
        }
        std::vector<float> packed(scene_packed_floats(scene));
        scene_pack(scene, packed.data());
        DevBuf<float> ds(packed.size()), dq((size_t)n * D);
        DevBuf<uint8_t> dout(n);
        PRRTC_HIP(hipMemcpy(ds.p, packed.data(), sizeof(float) * packed.size(), hipMemcpyHostToDevice));
        PRRTC_HIP(hipMemcpy(dq.p, poses, sizeof(float) * n * D, hipMemcpyHostToDevice));
        const size_t f4 = (packed.size() - 8) / 4;
        if (!scene_fits(f4)) {
            set_error("prrtc_check_poses: scene tables exceed %zu bytes", kMaxLdsScene);
            return PRRTC_EINVAL;
        }
        PRRTC_DISPATCH(robot, launch_check, ds.p, f4, dq.p, n, dout.p, (hipStream_t)0);
        PRRTC_HIP(hipMemcpy(out_free, dout.p, n, hipMemcpyDeviceToHost));
        return PRRTC_OK;
    });
}

int prrtc_validate_edges(int robot, const prrtc_scene *scene, const float *edges, int n, int granularity,
                         uint8_t *out_free) {
    return guarded("prrtc_validate_edges", [&]() -> int {
        const int D = robot_dof(robot);
        if (D < 0 || n < 0 || granularity < 1 || !scene_valid(scene) || (n > 0 && (!edges || !out_free))) {
            set_error("prrtc_validate_edges: invalid argument");
            return PRRTC_EINVAL;
        }
        if (n == 0) return PRRTC_OK;
        if (!all_finite(edges, (size_t)n * 2 * D)) {
            set_error("prrtc_validate_edges: non-finite edge endpoint");
            return PRRTC_EINVAL;
        }
        std::vector<float> packed(scene_packed_floats(scene));
        scene_pack(scene, packed.data());
        DevBuf<float> ds(packed.size()), de((size_t)n * 2 * D);
        DevBuf<uint8_t> dout(n);
        PRRTC_HIP(hipMemcpy(ds.p, packed.data(), sizeof(float) * packed.size(), hipMemcpyHostToDevice));
        PRRTC_HIP(hipMemcpy(de.p, edges, sizeof(float) * (size_t)n * 2 * D, hipMemcpyHostToDevice));
        const size_t f4 = (packed.size() - 8) / 4;
        if (!scene_fits(f4)) {
            set_error("prrtc_validate_edges: scene tables exceed %zu bytes", kMaxLdsScene);
            return PRRTC_EINVAL;
        }
        PRRTC_DISPATCH(robot, launch_validate, ds.p, f4, de.p, n, granularity, dout.p, (hipStream_t)0);
        PRRTC_HIP(hipMemcpy(out_free, dout.p, n, hipMemcpyDeviceToHost));
        return PRRTC_OK;
    });
}

int prrtc_validate_edges_async(int robot, const float *d_scene, size_t scene_floats, const float *d_edges, int n,
                               int granularity, uint8_t *d_out_free, void *stream) {
    return guarded("prrtc_validate_edges_async", [&]() -> int {
        const int D = robot_dof(robot);
        if (D < 0 || n < 0 || granularity < 1 || !d_scene || scene_floats < 8 || (scene_floats - 8) % 4 ||
            (n > 0 && (!d_edges || !d_out_free))) {
            set_error("prrtc_validate_edges_async: invalid argument");
            return PRRTC_EINVAL;
        }
        if (n == 0) return PRRTC_OK;
        if (!scene_fits((scene_floats - 8) / 4)) {
            set_error("prrtc_validate_edges_async: scene tables exceed %zu bytes", kMaxLdsScene);
            return PRRTC_EINVAL;
        }
        PRRTC_DISPATCH(robot, launch_validate, d_scene, (scene_floats - 8) / 4, d_edges, n, granularity, d_out_free,
                       (hipStream_t)stream);
        return PRRTC_OK;
    });
}

}  // extern "C"
