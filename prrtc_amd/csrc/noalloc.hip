// prrtc_solve without allocation: the planner's device state in a static
// device array.
//
// prrtc_solve keeps a per-thread workspace that it grows with hipMalloc and
// stages uploads through pinned host memory.  Callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) call prrtc_solve_noalloc instead: the problem's block (packed
// scene, start and goal rows, trees, parents, radii, Halton states, slot
// outcomes, staged chains), its PlanDesc / PlanState / PlanControl, the
// problem masks and the path rows all live in `scratch`, one static
// __device__ array per GPU, held until the process ends.  Uploads copy from
// the caller's arrays and from a per-thread pack buffer of static storage;
// the call runs on the null stream and returns when the solve is done.  The
// kernels and the launch shape are prrtc_solve's, so the results are the
// same bit for bit.
//
// Every call on a GPU uses that GPU's one array, which keeps the last
// call's contents: calls that overlap on one GPU race (one at a time per
// GPU).  Everything a solve reads is written by the call itself before its
// kernels run (uploads, then the init kernel's tree roots and Halton states).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "planner_host.h"

namespace prrtc {
namespace {

constexpr size_t kScratchBytes = PRRTC_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) char scratch[kScratchBytes];

// Device address of `scratch` on each GPU, taken once per device with that
// device current (the array's name is only a host-side stand-in).
char *g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

char *scratch_on(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw HipError{e != hipSuccess ? e : hipErrorInvalidSymbol,
                                                  "hipGetSymbolAddress(scratch)"};
        g_scratch[device] = static_cast<char *>(p);
    }
    return g_scratch[device];
}

struct Offsets {
    size_t off = 0;
    size_t take(size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    }
};

// Reverse rows [a, b) of a row-major array of dof-float rows, in place.
void reverse_rows(float *p, int a, int b, int dof) {
    for (int i = a, j = b - 1; i < j; i++, j--)
        for (int k = 0; k < dof; k++) std::swap(p[(size_t)i * dof + k], p[(size_t)j * dof + k]);
}

int solve_noalloc(const HostProblem &hp, prrtc_result *out, float *path, int path_cap) {
    const int64_t t0 = now_ns();
    if (int rc = check_problem(hp, "prrtc_solve_noalloc")) return rc;
    if (!out || !path || path_cap < 2) {
        set_error("prrtc_solve_noalloc: result/path buffers missing or path_cap < 2");
        return PRRTC_EINVAL;
    }
    const int D = robot_dof(hp.robot);
    const BlockLayout L = block_layout(hp);
    Offsets o;
    const size_t o_desc = o.take(sizeof(PlanDesc));
    const size_t o_state = o.take(sizeof(PlanState));
    const size_t o_ctrl = o.take(sizeof(PlanControl));
    const size_t o_mask = o.take(2 * sizeof(unsigned long long));   // running | open, one problem
    const size_t o_block = o.take(L.bytes);
    const size_t path_rows = (size_t)std::max(0, path_cap - 2);
    const size_t o_paths = o.take(sizeof(float) * path_rows * D);
    if (o.off > kScratchBytes) {
        set_error("prrtc_solve_noalloc: the problem needs %zu bytes of device state (trees of %d nodes, %zu path "
                  "rows); the static scratch holds %zu: lower max_samples or path_cap",
                  o.off, L.cap, path_rows, kScratchBytes);
        return PRRTC_ENOMEM;
    }
    int device = 0;
    PRRTC_HIP(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        set_error("prrtc_solve_noalloc: device %d beyond the %d supported", device, kMaxDevices);
        return PRRTC_EINVAL;
    }
    char *dev = scratch_on(device);
    const hipStream_t s = nullptr;   // the null stream: nothing to create

    // packed scene (static per-thread storage) and the caller's start and
    // goal rows, straight into the block
    static thread_local float packed[kMaxLdsScene / sizeof(float) + 8];
    scene_pack(hp.scene, packed);
    char *blk = dev + o_block;
    PRRTC_HIP(hipMemcpyAsync(blk + L.scene, packed, sizeof(float) * (size_t)L.scene_floats, hipMemcpyHostToDevice, s));
    PRRTC_HIP(hipMemcpyAsync(blk + L.init, hp.start, sizeof(float) * D, hipMemcpyHostToDevice, s));
    PRRTC_HIP(hipMemcpyAsync(blk + L.init + sizeof(float) * D, hp.goals, sizeof(float) * (size_t)hp.n_goals * D,
                             hipMemcpyHostToDevice, s));
    PlanDesc desc;
    fill_desc(desc, hp, L, blk);
    PlanControl ctrl{};
    ctrl.n_problems = 1;
    ctrl.mask_words = 1;
    ctrl.run_mask = reinterpret_cast<unsigned long long *>(dev + o_mask);
    ctrl.open_mask = ctrl.run_mask + 1;
    ctrl.path_cap = (int)std::min<size_t>(path_rows, INT32_MAX);
    ctrl.max_nnc = hp.st->num_new_configs;
    ctrl.prio_policy = kPrioOldest;
    ctrl.wide_slots = plan_wide_slots();
    ctrl.crowd_above = 0;
    const unsigned long long masks[2] = {1ull, 1ull};   // problem 0 runs, round 1 open
    PRRTC_HIP(hipMemcpyAsync(dev + o_desc, &desc, sizeof desc, hipMemcpyHostToDevice, s));
    PRRTC_HIP(hipMemcpyAsync(dev + o_ctrl, &ctrl, sizeof ctrl, hipMemcpyHostToDevice, s));
    PRRTC_HIP(hipMemcpyAsync(dev + o_mask, masks, sizeof masks, hipMemcpyHostToDevice, s));

    // prrtc_solve's kernels and launch shape for one problem
    const PlanDesc *d_desc = reinterpret_cast<const PlanDesc *>(dev + o_desc);
    PlanState *d_state = reinterpret_cast<PlanState *>(dev + o_state);
    PlanControl *d_ctrl = reinterpret_cast<PlanControl *>(dev + o_ctrl);
    float *d_paths = reinterpret_cast<float *>(dev + o_paths);
    const int waves = plan_task_waves(hp.robot);
    const int64_t tk = now_ns();
    launch_plan_init(hp.robot, d_desc, d_state, nullptr, 1, s);
    launch_plan_slice(hp.robot, d_desc, d_state, d_ctrl, d_paths, plan_grid(hp.robot, waves, hp.st->num_new_configs),
                      waves, (L.scene_floats - 8) / 4, L.max_ext, s);
    PlanState ps;
    PRRTC_HIP(hipMemcpyAsync(&ps, d_state, sizeof ps, hipMemcpyDeviceToHost, s));
    PRRTC_HIP(hipStreamSynchronize(s));
    const int64_t kernel_ns = now_ns() - tk;

    prrtc_result &r = *out;
    r = prrtc_result{};
    r.start_tree_size = ps.size[0];
    r.goal_tree_size = ps.size[1];
    r.edge_checks = (int64_t)ps.edge_checks;
    r.pose_checks = (int64_t)ps.pose_checks;
    r.solved = ps.status == kStatusSolved;
    r.status = r.solved ? PRRTC_OK : (ps.status == kStatusPathLimit ? PRRTC_EOVERFLOW : PRRTC_OK);
    r.iters = r.solved ? ps.iters : 0;     // as prrtc_solve reports them, also when the path overflows
    r.cost = r.solved ? ps.cost : 0.0f;
    if (r.solved) {
        // goal -> start (pRRTC.cu:715-746): the goal reached, the tree-1
        // trace reversed, the tree-0 trace, the start.  The two traces are
        // copied behind row 0 as the commit wrote them (t-trace, o-trace)
        // and rearranged in place.
        const int n_rows = ps.t_len + ps.o_len + 2;
        if (n_rows > path_cap) {
            r.solved = 0;
            r.status = PRRTC_EOVERFLOW;
        } else {
            const int tl = ps.t_len, ol = ps.o_len, used = tl + ol;
            if (used > 0)
                PRRTC_HIP(hipMemcpy(path + D, d_paths + (size_t)ps.path_off * D, sizeof(float) * (size_t)used * D,
                                    hipMemcpyDeviceToHost));
            if (ps.t_tree == 1) {
                reverse_rows(path, 1, 1 + tl, D);            // [t reversed, o]
            } else {
                reverse_rows(path, 1, 1 + used, D);          // [o reversed, t reversed]
                reverse_rows(path, 1 + ol, 1 + used, D);     // [o reversed, t]
            }
            std::copy_n(hp.goals + (size_t)ps.reached * D, D, path);
            std::copy_n(hp.start, D, path + (size_t)(n_rows - 1) * D);
            r.path_nodes = n_rows;
            r.path_length = used;
        }
    }
    r.kernel_ns = kernel_ns;
    r.wall_ns = now_ns() - t0;
    return PRRTC_OK;
}

}  // namespace
}  // namespace prrtc

using namespace prrtc;

extern "C" int prrtc_solve_noalloc(int robot, const float *start, const float *goals, int n_goals,
                                   const prrtc_scene *scene, const prrtc_settings *settings, prrtc_result *result,
                                   float *path, int path_cap) {
    return guarded("prrtc_solve_noalloc", [&]() -> int {
        if (!result) {
            set_error("prrtc_solve_noalloc: result is NULL");
            return PRRTC_EINVAL;
        }
        *result = prrtc_result{};
        return solve_noalloc(HostProblem{robot, start, goals, n_goals, scene, settings, 0}, result, path, path_cap);
    });
}
