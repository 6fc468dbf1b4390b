// Device building blocks of the pRRTC hot path for gfx950 (CDNA4, wave64).
//
//  * packed scene layout (float4 tables) and its LDS staging;
//  * sphere-vs-obstacle tests in the reference's squared form
//    (src/planning/utils.cuh:58-208): value < 0 means collision;
//  * check contexts that the generated robot code drives, with the
//    wavefront-level early exit (ballot after every check group);
//  * robot dispatch (Sphere / Panda / Fetch / Baxter).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "robots/baxter.gen.h"
#include "robots/fetch.gen.h"
#include "robots/panda.gen.h"
#include "robots/ur5.gen.h"

namespace prrtc {

constexpr int kWave = 64;
constexpr int kXcds = 8;   // MI355X: 8 XCDs, workgroups dispatched to them round-robin
constexpr int kMaxDof = 14;
// Largest obstacle table set staged into LDS per workgroup.
constexpr size_t kMaxLdsScene = 64 * 1024;

// ---------------------------------------------------------------------------
// packed scene: a header of two float4 (int counts), then the tables
//   spheres  : 1 float4 each   (x, y, z, r)
//   capsules : 2 float4 each   (x1, y1, z1, r) (xv, yv, zv, rdv)
//   z-caps   : 2 float4 each
//   cuboids  : 4 float4 each   (x, y, z, r1) (a1, r2) (a2, r3) (a3, cull)
//   z-cuboids: 4 float4 each   (cull: padded half-diagonal, set by the host pack)
struct SceneHeader {
    int32_t n_sph, n_cap, n_zcap, n_cub, n_zcub, n_f4, pad0, pad1;
};
static_assert(sizeof(SceneHeader) == 32, "header is two float4");

__host__ __device__ inline int scene_payload_f4(int n_sph, int n_cap, int n_zcap, int n_cub, int n_zcub) {
    return n_sph + 2 * (n_cap + n_zcap) + 4 * (n_cub + n_zcub);
}

// Obstacle tables as seen by one workgroup: counts in registers (uniform),
// payload behind a pointer (LDS after staging, or global).
struct Obstacles {
    const float4 *sph, *cap, *zcap, *cub, *zcub;
    int n_sph, n_cap, n_zcap, n_cub, n_zcub;

    __device__ void bind(const SceneHeader &h, const float4 *payload) {
        n_sph = h.n_sph; n_cap = h.n_cap; n_zcap = h.n_zcap; n_cub = h.n_cub; n_zcub = h.n_zcub;
        sph = payload;
        cap = sph + n_sph;
        zcap = cap + 2 * n_cap;
        cub = zcap + 2 * n_zcap;
        zcub = cub + 4 * n_cub;
    }

    // Share `part` of 2^shift: a contiguous share of every table (the shares
    // of one table partition it), cut by shifts (cheap enough to take at every
    // obstacle loop instead of holding a sliced copy in registers).
    __device__ __forceinline__ Obstacles slice_pow2(int part, int shift) const {
        Obstacles o = *this;
        cut2(o.sph, o.n_sph, 1, part, shift);
        cut2(o.cap, o.n_cap, 2, part, shift);
        cut2(o.zcap, o.n_zcap, 2, part, shift);
        cut2(o.cub, o.n_cub, 4, part, shift);
        cut2(o.zcub, o.n_zcub, 4, part, shift);
        return o;
    }

private:
    __device__ static void cut2(const float4 *&p, int &n, int stride, int part, int shift) {
        const int lo = (n * part) >> shift, hi = (n * (part + 1)) >> shift;
        p += (size_t)lo * stride;
        n = hi - lo;
    }
};

// Read a packed scene's header.  `cap_f4` is the payload size the host
// allocated; a header that does not fit it (or is inconsistent) is read as an
// empty scene, so no access leaves the buffer.
__device__ inline void read_header(const float *packed, int cap_f4, SceneHeader &h) {
    const int4 *hdr = reinterpret_cast<const int4 *>(packed);
    const int4 h0 = hdr[0], h1 = hdr[1];
    h.n_sph = h0.x; h.n_cap = h0.y; h.n_zcap = h0.z; h.n_cub = h0.w;
    h.n_zcub = h1.x; h.n_f4 = h1.y; h.pad0 = 0; h.pad1 = 0;
    const bool ok = h.n_sph >= 0 && h.n_cap >= 0 && h.n_zcap >= 0 && h.n_cub >= 0 && h.n_zcub >= 0 &&
                    h.n_f4 == scene_payload_f4(h.n_sph, h.n_cap, h.n_zcap, h.n_cub, h.n_zcub) && h.n_f4 <= cap_f4;
    if (!ok) h = SceneHeader{0, 0, 0, 0, 0, 0, 0, 0};
}

// Copy a packed scene's payload into LDS; every lane of the workgroup helps.
__device__ inline void stage_scene(const float *packed, int cap_f4, float4 *lds, SceneHeader &h) {
    read_header(packed, cap_f4, h);
    const float4 *src = reinterpret_cast<const float4 *>(packed) + 2;
    for (int i = threadIdx.x; i < h.n_f4; i += blockDim.x) lds[i] = src[i];
    __syncthreads();
}

// Bind a packed scene for a workgroup: staged into LDS, or read in place.
template <bool LDS>
__device__ inline void bind_scene(const float *packed, int cap_f4, float4 *lds, Obstacles &obs) {
    SceneHeader h;
    if (LDS) {
        stage_scene(packed, cap_f4, lds, h);
        obs.bind(h, lds);
    } else {
        read_header(packed, cap_f4, h);
        obs.bind(h, reinterpret_cast<const float4 *>(packed) + 2);
    }
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// sphere_environment_in_collision (utils.cuh:211-242) for N spheres per lane
// at once: OR over the five tables, each obstacle read once (an LDS
// broadcast: every lane reads the same address) and tested against all N
// spheres.  Tests are the reference's squared forms (utils.cuh:58-208),
// value < 0 meaning collision.  With EARLY the wavefront leaves the loops as
// soon as every active lane has a hit in one of its N spheres.
//
// Rounding is pinned: every multiply-add below is an explicit fmaf in a fixed
// association (and contraction is off), so the CPU oracle can reproduce each
// verdict bit for bit in its "fused" mode.
__device__ __forceinline__ float sq3(float dx, float dy, float dz) { return fmaf(dx, dx, fmaf(dy, dy, dz * dz)); }

// Obstacle table read.  LOCAL: the tables are known to be staged in LDS, so
// the generic pointer is narrowed to an LDS address (its low 32 bits) and
// read with an LDS instruction, where the compiler cannot see the pointer's
// origin (a caller that is not inlined into the kernel) and would otherwise
// emit a flat load.
template <bool LOCAL>
__device__ __forceinline__ float4 ld_obs(const float4 *p) {
    if constexpr (LOCAL) {
        typedef float f4v __attribute__((ext_vector_type(4)));
        typedef const __attribute__((address_space(3))) f4v local_f4v;
        const f4v v = *reinterpret_cast<local_f4v *>((uint32_t)(uintptr_t)p);
        return make_float4(v.x, v.y, v.z, v.w);
    } else {
        return *p;
    }
}

// GROUP: the N spheres all lie inside the ball (gx, gy, gz; gr) on every
// lane -- the leaves of one link, bounded about the link's gate centre with a
// 1 mm pad by the generator.  An obstacle whose own bounding ball (sphere:
// itself; cuboid, capsule: the cull balls below) stays outside that ball on
// every lane is at least 1 mm from every one of the N spheres, so all N of
// its tests are positive and the wave skips the obstacle with one test.
//
// ANY: only whether some of the N spheres hits is wanted (a leaf group, whose
// hits the caller ORs): one running minimum over all N spheres replaces the N
// per-sphere minima (some value < 0 iff their minimum is), which frees N - 1
// registers of every leaf group; hit[0] carries the answer.
template <int N, bool EARLY, bool LOCAL = false, bool GROUP = false, bool ANY = false>
__device__ __forceinline__ void env_multi(const Obstacles &o, const float (&x)[N], const float (&y)[N],
                                          const float (&z)[N], const float (&r)[N], bool (&hit)[ANY ? 1 : N],
                                          float gx = 0.0f, float gy = 0.0f, float gz = 0.0f, float gr = 0.0f) {
#pragma clang fp contract(off)
    auto group_far = [&](float cx, float cy, float cz, float cr) {
        const float rb = gr + cr;
        return ballot(!(sq3(cx - gx, cy - gy, cz - gz) > rb * rb)) == 0;
    };
    // Each sphere keeps the smallest test value over the obstacles seen: some
    // value < 0 (a collision) iff the minimum is < 0.  One v_min per test
    // replaces a compare, a select and the bit packing of a boolean OR; a
    // NaN test value is dropped by fminf as it would fail `< 0`.
    constexpr int M = ANY ? 1 : N;   // running minima kept
    float m[M];
#pragma unroll
    for (int n = 0; n < M; n++) m[n] = __builtin_huge_valf();
    auto done = [&]() {
        bool any_hit = false;
#pragma unroll
        for (int n = 0; n < M; n++) any_hit |= m[n] < 0.0f;
        return ballot(!any_hit) == 0;
    };
    auto finish = [&]() {
#pragma unroll
        for (int n = 0; n < M; n++) hit[n] = m[n] < 0.0f;
    };
    for (int i = 0; i < o.n_sph; i++) {                       // utils.cuh:58-81
        const float4 s = ld_obs<LOCAL>(o.sph + i);
        if (GROUP && group_far(s.x, s.y, s.z, s.w)) continue;
#pragma unroll
        for (int n = 0; n < N; n++) {
            const float rs = s.w + r[n];
            m[ANY ? 0 : n] = fminf(m[ANY ? 0 : n], fmaf(-rs, rs, sq3(s.x - x[n], s.y - y[n], s.z - z[n])));
        }
        if (EARLY && (i & 7) == 7 && done()) return finish();
    }
    if (EARLY && o.n_sph && done()) return finish();
    // Capsules are culled like cuboids (below) against the ball about the
    // segment's midpoint that holds the capsule: radius |v|/2 + r, padded by
    // 2^-16 relative and 1 mm absolute.  The pad covers the rounding of
    // sqrtf and of both tests many times over.
    for (int i = 0; i < o.n_cap; i++) {                       // utils.cuh:113-126
        const float4 p = ld_obs<LOCAL>(o.cap + 2 * i), v = ld_obs<LOCAL>(o.cap + 2 * i + 1);
        const float mx = fmaf(0.5f, v.x, p.x), my = fmaf(0.5f, v.y, p.y), mz = fmaf(0.5f, v.z, p.z);
        const float cull = fmaf(0.5f * sqrtf(sq3(v.x, v.y, v.z)) + p.w, 1.0f + 0x1p-16f, 1e-3f);
        if (GROUP && group_far(mx, my, mz, cull)) continue;
#pragma unroll
        for (int n = 0; n < N; n++) {
            const float rb = r[n] + cull;
            if (ballot(!(sq3(x[n] - mx, y[n] - my, z[n] - mz) > rb * rb)) == 0) continue;
            const float dot = fmaf(x[n] - p.x, v.x, fmaf(y[n] - p.y, v.y, (z[n] - p.z) * v.z));
            const float cdf = fmaxf(fminf(dot * v.w, 1.0f), 0.0f);
            const float rs = r[n] + p.w;
            const float d = sq3(x[n] - fmaf(v.x, cdf, p.x), y[n] - fmaf(v.y, cdf, p.y), z[n] - fmaf(v.z, cdf, p.z));
            m[ANY ? 0 : n] = fminf(m[ANY ? 0 : n], fmaf(-rs, rs, d));
        }
        if (EARLY && (i & 7) == 7 && done()) return finish();
    }
    if (EARLY && o.n_cap && done()) return finish();
    for (int i = 0; i < o.n_zcap; i++) {                      // utils.cuh:135-148
        const float4 p = ld_obs<LOCAL>(o.zcap + 2 * i), v = ld_obs<LOCAL>(o.zcap + 2 * i + 1);
#pragma unroll
        for (int n = 0; n < N; n++) {
            const float dot = (z[n] - p.z) * v.z;
            const float cdf = fmaxf(fminf(dot * v.w, 1.0f), 0.0f);
            const float rs = r[n] + p.w;
            const float d = sq3(x[n] - p.x, y[n] - p.y, z[n] - fmaf(v.z, cdf, p.z));
            m[ANY ? 0 : n] = fminf(m[ANY ? 0 : n], fmaf(-rs, rs, d));
        }
        if (EARLY && (i & 7) == 7 && done()) return finish();
    }
    if (EARLY && o.n_zcap && done()) return finish();
    // Cuboids carry a cull radius (a3.w): the box's half-diagonal with a 1 mm
    // pad, from the host pack.  When no lane's sphere comes within r + cull
    // of the centre, the box test is positive by at least that pad on every
    // lane and cannot lower m below 0, so the wave skips it (the test itself,
    // where it runs, is unchanged).
    for (int i = 0; i < o.n_cub; i++) {                       // utils.cuh:158-175
        const float4 c = ld_obs<LOCAL>(o.cub + 4 * i), a1 = ld_obs<LOCAL>(o.cub + 4 * i + 1),
                     a2 = ld_obs<LOCAL>(o.cub + 4 * i + 2), a3 = ld_obs<LOCAL>(o.cub + 4 * i + 3);
        if (GROUP && group_far(c.x, c.y, c.z, a3.w)) continue;
#pragma unroll
        for (int n = 0; n < N; n++) {
            const float xs = x[n] - c.x, ys = y[n] - c.y, zs = z[n] - c.z;
            const float rb = r[n] + a3.w;
            if (ballot(!(sq3(xs, ys, zs) > rb * rb)) == 0) continue;
            const float d1 = fmaxf(0.0f, fabsf(fmaf(a1.x, xs, fmaf(a1.y, ys, a1.z * zs))) - c.w);
            const float d2 = fmaxf(0.0f, fabsf(fmaf(a2.x, xs, fmaf(a2.y, ys, a2.z * zs))) - a1.w);
            const float d3 = fmaxf(0.0f, fabsf(fmaf(a3.x, xs, fmaf(a3.y, ys, a3.z * zs))) - a2.w);
            m[ANY ? 0 : n] = fminf(m[ANY ? 0 : n], fmaf(-r[n], r[n], sq3(d1, d2, d3)));
        }
        if (EARLY && (i & 7) == 7 && done()) return finish();
    }
    if (EARLY && o.n_cub && done()) return finish();
    for (int i = 0; i < o.n_zcub; i++) {                      // utils.cuh:184-201
        const float4 c = ld_obs<LOCAL>(o.zcub + 4 * i), a1 = ld_obs<LOCAL>(o.zcub + 4 * i + 1),
                     a2 = ld_obs<LOCAL>(o.zcub + 4 * i + 2);
        const float cull = ld_obs<LOCAL>(o.zcub + 4 * i + 3).w;
        if (GROUP && group_far(c.x, c.y, c.z, cull)) continue;
#pragma unroll
        for (int n = 0; n < N; n++) {
            const float xs = x[n] - c.x, ys = y[n] - c.y, zs = z[n] - c.z;
            const float rb = r[n] + cull;
            if (ballot(!(sq3(xs, ys, zs) > rb * rb)) == 0) continue;
            const float d1 = fmaxf(0.0f, fabsf(fmaf(a1.x, xs, a1.y * ys)) - c.w);
            const float d2 = fmaxf(0.0f, fabsf(fmaf(a2.x, xs, a2.y * ys)) - a1.w);
            const float d3 = fmaxf(0.0f, fabsf(zs) - a2.w);
            m[ANY ? 0 : n] = fminf(m[ANY ? 0 : n], fmaf(-r[n], r[n], sq3(d1, d2, d3)));
        }
        if (EARLY && (i & 7) == 7 && done()) return finish();
    }
    finish();
}

// sphere_sphere_self_collision (utils.cuh:83-86) with the radius sum folded.
__device__ __forceinline__ bool self_hit(float ax, float ay, float az, float bx, float by, float bz, float rs) {
#pragma clang fp contract(off)
    return fmaf(-rs, rs, sq3(ax - bx, ay - by, az - bz)) < 0.0f;
}

// ---------------------------------------------------------------------------
// Check context driven by the generated robot code, for a pass of up to 64
// poses (one per lane).
//
// Path mode (independent == false): `key` orders the lanes' poses into groups
// (sub-edges, non-decreasing with the lane index).  Lanes whose key is >= the
// key of the lowest colliding lane can no longer change the answer, so they
// stop; the wavefront leaves once no lane is live.  With one key for the
// whole wave (one edge) the first collision anywhere ends the pass.
//
// Pose mode (independent == true): every lane is its own question; a lane
// stops only on its own collision.
struct CheckCtx {
    Obstacles obs;
    bool coll;
    bool dead;
    bool independent;
    int key;

    __device__ __forceinline__ bool live() const { return !coll && !dead; }
    __device__ __forceinline__ bool any(bool p) const { return ballot(p) != 0; }
    // Self-collision group j (numbered in check order) is this wave's to
    // test; a lone wave tests every group.
    __device__ __forceinline__ bool mine(int) const { return true; }
    template <int N, bool EARLY>
    __device__ __forceinline__ void env_multi(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                              const float (&r)[N], bool (&hit)[N]) const {
        prrtc::env_multi<N, EARLY>(obs, x, y, z, r, hit);
    }
    // Whether any of the N spheres hits an obstacle (a leaf group).
    template <int N, bool EARLY>
    __device__ __forceinline__ bool env_any(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                            const float (&r)[N]) const {
        bool h[1];
        prrtc::env_multi<N, EARLY, false, false, true>(obs, x, y, z, r, h);
        return h[0];
    }
    // The same for N spheres that lie inside the ball (gx, gy, gz; gr) on
    // every lane: an obstacle whose bounding ball stays outside it is skipped
    // for the whole group (GROUP).  With one running minimum per group the
    // bound fits the one-wave-per-edge kernels' register budget (Panda 209
    // VGPRs, Fetch 255: two waves per SIMD, no scratch).
    template <int N, bool EARLY>
    __device__ __forceinline__ bool env_any_bounded(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                                    const float (&r)[N], float gx, float gy, float gz,
                                                    float gr) const {
        bool h[1];
        prrtc::env_multi<N, EARLY, false, true, true>(obs, x, y, z, r, h, gx, gy, gz, gr);
        return h[0];
    }
    __device__ __forceinline__ bool self(float ax, float ay, float az, float bx, float by, float bz, float rs) const {
        return self_hit(ax, ay, az, bx, by, bz, rs);
    }
    __device__ __forceinline__ void hit() { coll = true; }
    // Gate bits of this lane's pose after the gate pass.  A context whose
    // waves split the obstacle tables ORs them across its waves; here every
    // wave sees every obstacle, so there is nothing to share.
    __device__ __forceinline__ void share_gates(uint64_t &) const {}
    __device__ __forceinline__ bool stop() {
        if (!independent) {
            const uint64_t cm = ballot(coll);
            if (cm) {
                const int l = __builtin_ctzll(cm);
                const int stop_key = __builtin_amdgcn_readlane(key, l);
                dead = dead || key >= stop_key;
            }
        }
        return ballot(!coll && !dead) == 0;
    }
};

// ---------------------------------------------------------------------------
// robots

struct SphereRobot {     // Robots.hh:190-237; fkcc<Sphere>, utils.cuh:330-334
    static constexpr int DOF = 3;
    static constexpr int NSPHERES = 1;
    __host__ __device__ static constexpr float lower(int) { return -5.0f; }
    __host__ __device__ static constexpr float span(int) { return 10.0f; }
    template <class Ctx>
    __device__ __forceinline__ static void check(const float *q, Ctx &ctx) {
        const float x[1] = {q[0]}, y[1] = {q[1]}, z[1] = {q[2]}, r[1] = {1.0f};
        bool h[1];
        if (ctx.live()) {
            ctx.template env_multi<1, false>(x, y, z, r, h);
            if (h[0]) ctx.hit();
        }
        ctx.stop();
    }
    __device__ __forceinline__ static void fk_spheres(const float *q, float *out) {
        out[0] = q[0]; out[1] = q[1]; out[2] = q[2];
    }
};

#define PRRTC_GEN_ROBOT(NAME, NS)                                                              \
    struct NAME##Robot {                                                                       \
        static constexpr int DOF = gen::NS::DOF;                                               \
        static constexpr int NSPHERES = gen::NS::NSPHERES;                                     \
        __host__ __device__ static constexpr float lower(int i) { return gen::NS::lower(i); } \
        __host__ __device__ static constexpr float span(int i) { return gen::NS::span(i); }   \
        template <class Ctx>                                                                   \
        __device__ __forceinline__ static void check(const float *q, Ctx &ctx) {               \
            gen::NS::check(q, ctx);                                                            \
        }                                                                                      \
        __device__ __forceinline__ static void fk_spheres(const float *q, float *out) {       \
            gen::NS::fk_spheres(q, out);                                                       \
        }                                                                                      \
    };
PRRTC_GEN_ROBOT(Panda, panda)
PRRTC_GEN_ROBOT(Fetch, fetch)
PRRTC_GEN_ROBOT(Baxter, baxter)
PRRTC_GEN_ROBOT(Ur5, ur5)
#undef PRRTC_GEN_ROBOT

// Every robot the library is built for, as X(id, struct): the one list the
// dispatch switches expand.  A robot from the model compiler is one more
// line here, its generated header above and its id in include/prrtc.h.
#define PRRTC_FOR_EACH_ROBOT(X)        \
    X(PRRTC_ROBOT_SPHERE, SphereRobot) \
    X(PRRTC_ROBOT_PANDA, PandaRobot)   \
    X(PRRTC_ROBOT_FETCH, FetchRobot)   \
    X(PRRTC_ROBOT_BAXTER, BaxterRobot) \
    X(PRRTC_ROBOT_UR5, Ur5Robot)
constexpr int kNumRobots = 5;   // ids 0 .. kNumRobots - 1
constexpr bool robot_ids_dense() {
    int n = 0, top = -1;
#define PRRTC_COUNT_ROBOT(ID, R) n++; top = ID > top ? ID : top;
    PRRTC_FOR_EACH_ROBOT(PRRTC_COUNT_ROBOT)
#undef PRRTC_COUNT_ROBOT
    return n == kNumRobots && top == kNumRobots - 1;
}
static_assert(robot_ids_dense(), "PRRTC_FOR_EACH_ROBOT lists ids 0 .. kNumRobots - 1");
// The type of a robot in a call expanded from PRRTC_FOR_EACH_ROBOT.
template <class R> struct RobotTag { using type = R; };

}  // namespace prrtc
