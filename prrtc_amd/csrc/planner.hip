// Round-synchronous persistent pRRTC planner (gfx950) and its C ABI:
// prrtc_solve (pRRTC::solve<Robot>, src/planning/Planners.hh:75-78) and
// prrtc_solve_batch (many problems per launch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "planner_host.h"

// Minimum waves per SIMD the persistent planner is compiled for (bounds its
// registers: 1 -> up to 512, 2 -> 256).
#ifndef PRRTC_PLAN_WAVES_PER_EU
#define PRRTC_PLAN_WAVES_PER_EU 1
#endif

namespace prrtc {

// ------------------------------------------------------------------ device

// Combine one value per wave across the workgroup: the OR (want_min false)
// or the minimum of `v`.  `buf` holds 2 * kMaxTaskWaves ints in LDS, used by
// alternate calls (`parity`), so one barrier per call suffices.
__device__ __forceinline__ int across_waves(int v, bool want_min, int *buf, int parity, const Team &tm) {
    const int W = tm.waves;
    if (W == 1) return v;
    int *b = buf + parity * kMaxTaskWaves;
    if (lane_id() == 0) b[tm.wave] = v;
    tm.sync();
    int r = b[0];
    for (int w = 1; w < W; w++) r = want_min ? min(r, b[w]) : (r | b[w]);
    return __builtin_amdgcn_readfirstlane(r);   // alike on every lane: keeps the pass loop's exits scalar
}

// Check context of a planner task.  The task's waves validate the same 64
// poses per pass and split the obstacle tables between them (each wave's
// `obs` is one slice), so a pass costs a fraction of one wave's work; a
// pose's gate bits are ORed across the waves before the leaf checks, and
// the waves' collision flags are combined after the pass.  A gate hits iff
// some obstacle hits it, whichever wave tests that obstacle, so the verdicts
// are those of a single wave that sees every obstacle.
//
// The waves also share their early exits: in each pass a wave that finds a
// collision posts its lowest colliding key to an LDS word (atomic min), and
// at every check-group boundary each wave retires its lanes whose key is at
// or above the posted one.  Such lanes cannot lower the pass's answer (the
// lowest colliding key over all waves, or "any collision" on one edge), so
// a collision found on any wave ends the pass on all of them without a
// barrier.  The word is chosen by pass parity and reset a pass ahead, so a
// wave never reads a previous pass's key.
//
// Self-collision groups, which need no obstacles, are dealt out instead:
// group j goes to wave j mod W (`mine`), and a hit on any wave ends the pass
// on all of them through the same word.  Every group is still tested by one
// wave for every live lane, so the verdicts are unchanged.
// The two words live in static LDS, so every access is a plain LDS
// instruction; only a task of the whole workgroup uses them.
__device__ __forceinline__ int *task_stop_words() {
    __shared__ int words[2];
    return words;
}

struct TaskCtx : CheckCtx {
    uint64_t *gbuf;   // LDS, kMaxTaskWaves * kWave words
    int pass;         // index of the current pass within the task
    int waves;        // waves sharing this pass's poses (1: the wave has poses of its own)
    int gbase;        // first wave of those
    // this wave's share of the obstacles: part `part` of 2^shift (obs holds
    // the whole staged scene; the share is cut at every obstacle loop)
    int part, shift;
    bool shared_stop; // the task runs on several waves: stop keys go through LDS
    bool wide_ok;     // its chain passes may give each wave poses of its own
    unsigned mine12;  // bit i: self groups j with j mod 12 == i are this wave's (12 = lcm of 1..4)
    __device__ __forceinline__ void share_gates(uint64_t &gate) {
        const int W = waves;
        if (W == 1) return;
        gbuf[threadIdx.x] = gate;
        __syncthreads();
        uint64_t g = 0;
        const int lane = lane_id();
        for (int w = 0; w < W; w++) g |= gbuf[(gbase + w) * kWave + lane];
        __syncthreads();   // gbuf is written again by the next pass
        gate = g;
    }
    static_assert(kMaxTaskWaves <= 4, "the 12-group pattern of mine12 needs 12 % waves == 0");
    __device__ __forceinline__ bool mine(int j) const { return (mine12 >> (j % 12)) & 1u; }
    // The obstacle tables of a task are always staged in LDS.
    template <int N, bool EARLY>
    __device__ __forceinline__ void env_multi(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                              const float (&r)[N], bool (&hit)[N]) const {
        prrtc::env_multi<N, EARLY, true>(obs.slice_pow2(part, shift), x, y, z, r, hit);
    }
    template <int N, bool EARLY>
    __device__ __forceinline__ bool env_any(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                            const float (&r)[N]) const {
        bool h[1];
        prrtc::env_multi<N, EARLY, true, false, true>(obs.slice_pow2(part, shift), x, y, z, r, h);
        return h[0];
    }
    template <int N, bool EARLY>
    __device__ __forceinline__ bool env_any_bounded(const float (&x)[N], const float (&y)[N], const float (&z)[N],
                                                    const float (&r)[N], float gx, float gy, float gz,
                                                    float gr) const {
        bool h[1];
        prrtc::env_multi<N, EARLY, true, true, true>(obs.slice_pow2(part, shift), x, y, z, r, h, gx, gy, gz, gr);
        return h[0];
    }
    // CheckCtx::stop in path mode (a planner task is never in pose mode),
    // with the stop key shared across waves.
    __device__ __forceinline__ bool stop() {
        const uint64_t cm = ballot(coll);
        int sk = cm ? __builtin_amdgcn_readlane(key, __builtin_ctzll(cm)) : 0x7fffffff;
        if (shared_stop) {
            int *w = task_stop_words() + (pass & 1);
            if (cm && lane_id() == 0) __hip_atomic_fetch_min(w, sk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            sk = min(sk, __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
        }
        dead = dead || key >= sk;
        return ballot(!coll && !dead) == 0;
    }
};

// One task: slot s of problem p in the current round (pRRTC.cu:342-551),
// run by the threads of `tm`: the whole workgroup (1..kMaxTaskWaves waves)
// or one wave on its own.
//
// The steered edge and the connect chain are validated by one loop with a
// single call site of the robot's check sequence (the sequence is large; one
// inlined copy keeps register use down); a pass puts 64 poses on the lanes
// of every wave, the waves splitting the obstacles (TaskCtx).  Uniform
// vectors live in LDS: chain[i] holds the chain configurations c_0 = new
// node, c_{i+1} = c_i + vec; uv holds the nearest node, the steered step and
// the connect step; red and wbuf are the cross-wave reduction buffers.
// Inlined into k_plan (its one caller): as a called function it saved and
// restored ~36 callee-saved VGPRs per task through the stack (144 B per
// lane), which wrote tens of GB per batch through L2; inlined, the planner
// kernels use no scratch at all.
template <class Robot>
__device__ __attribute__((always_inline)) void plan_task(const PlanDesc &d, PlanState *st, int s, const TaskCtx &ctx_in, float *chain, float *uv,
                          float *red, int *wbuf, const Team &tm) {
#pragma clang fp contract(off)
    // a private copy: its fields stay in registers across the check sequence
    // (through the caller's reference they would be reloaded after every
    // store to the pass flags)
    TaskCtx ctx = ctx_in;   // ctx.obs: the whole staged scene
    ctx.shared_stop = tm.waves > 1;
    // A pass puts 64 poses on each of G pose groups of the task's waves (G =
    // 1, 2 or all of them).  A group's waves test the same poses and split
    // the obstacles and self groups between them, which ends short work
    // soonest (one edge, a chain's first sub-edges); a group of one wave
    // tests its poses against the whole scene, which computes no check twice
    // and needs no gate exchange (long chains).  Every pose gets the full
    // check whatever the shape, and a collision on any wave retires the
    // lanes at or above its key on all of them.
    int groups = 0;   // G of the current shape
    int gw = 0;       // this wave's pose group
    // (task waves are 1, 2 or 4, so a group is 2^k waves)
    auto shape = [&](int G) {
        // G derives from values every lane holds alike (the chain length),
        // which the compiler cannot tell: made wave-uniform here, the share
        // stays scalar, and with it the obstacle loops and group branches
        G = __builtin_amdgcn_readfirstlane(G);
        if (G == groups) return;
        groups = G;
        const int gs = tm.waves / G;          // waves per group
        gw = tm.wave / gs;
        const int wi = tm.wave - gw * gs;     // this wave's index in its group
        ctx.part = wi;
        ctx.shift = gs == 4 ? 2 : gs == 2 ? 1 : 0;
        ctx.waves = gs;
        ctx.gbase = gw * gs;
        unsigned m = 0;
        for (int i = 0; i < 12; i++) m |= (unsigned)(i % gs == wi) << i;
        // uniform over the wave: the group branches stay scalar branches
        ctx.mine12 = (unsigned)__builtin_amdgcn_readfirstlane((int)m);
    };
    const bool wide_ok = __builtin_amdgcn_readfirstlane((int)ctx_in.wide_ok) != 0 && d.wide_from >= 0;
    // pose groups for n poses: one pass where the waves allow it
    auto groups_for = [&](int n) {
        const int want = (n + kWave - 1) / kWave;
        if (want <= 1 || tm.waves == 1) return 1;
        if (want == 2 && (tm.waves & 1) == 0) return 2;
        return tm.waves;
    };
    constexpr int D = Robot::DOF;
    const int tid = tm.tid, NT = tm.nt;
    const int lane = lane_id();
    const int iter = st->round;
    const int n0 = st->size[0], n1 = st->size[1];
    // the pass loop's bounds, wave-uniform to the compiler too: the group
    // shares and the ctx fields carried through the loop stay scalar
    const int g = __builtin_amdgcn_readfirstlane(d.g);
    const float gf = (float)g;
    int *meta = d.meta + (size_t)s * 8;
    float *stage = d.stage + (size_t)s * (d.max_ext + 1) * D;
    float *nnv = uv, *delta = uv + D, *step = uv + 2 * D, *vecv = uv + 3 * D;
    int edges = 0;
    int *stopw = task_stop_words();
    if (tid == 0 && tm.waves > 1) stopw[0] = stopw[1] = 0x7fffffff;   // read only after the next barrier

    // tree choice and Halton sample (uniform across the workgroup)
    const int t = choose_tree(d, s, iter, d.tprev[s], n0, n1);
    const int o = 1 - t;
    float *hs = d.halton + (size_t)s * 3 * D;
    float sample[D], h_n[D], h_d[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        float b = hs[k], n = hs[D + k], dd = hs[2 * D + k];
        const float u = halton_step(b, n, dd);
        h_n[k] = n;
        h_d[k] = dd;
        sample[k] = fmaf(u, Robot::span(k), Robot::lower(k));     // Robots.hh scale_cfg
    }
    // every wave has read the slot's Halton state and tree before thread 0
    // advances them
    if (tm.waves > 1) tm.sync();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < D; k++) { hs[D + k] = h_n[k]; hs[2 * D + k] = h_d[k]; }
        d.tprev[s] = t;
    }

    // nearest node in the target tree (pRRTC.cu:384-418)
    const float *tn = d.nodes + (size_t)t * D * d.cap;
    float bsq;
    const int nn = nearest<D>(tn, d.cap, t == 0 ? n0 : n1, sample, bsq, red, tm);
    const float dist = sqrtf(bsq);
    const float scale = fminf(1.0f, d.range / dist);
    int flags = 0, n_ok = 0, nn_o = 0;
    int poses = 0;   // poses put on lanes (the same on every wave of the team)
    if (d.dd && d.radii[(size_t)t * d.cap + nn] < dist) {          // pRRTC.cu:420-427
        flags = kSkip;
    } else {
        if (tid < D) {                                                // pRRTC.cu:430-434
            const float a = tn[(size_t)tid * d.cap + nn];
            float sl = sample[0];
#pragma unroll
            for (int k = 1; k < D; k++) if (tid == k) sl = sample[k];
            const float c = fmaf(sl - a, scale, a);
            nnv[tid] = a;
            chain[tid] = c;
            delta[tid] = (c - a) / gf;
        }
        tm.sync();
        // phase 0: the steered edge, poses nn + (k+1) delta in interleaved passes;
        // phase 1: the connect chain, pose j on sub-edge j / g at c_i + (j%g + 1) step
        const int G0 = __builtin_amdgcn_readfirstlane(d.wide_edge ? groups_for(g) : 1);
        const int passes = (g + kWave * G0 - 1) / (kWave * G0);
        shape(G0);
        int phase = 0, p = 0, base = 0, total = 0, n_chain = 0, npass = 0, cpass = 0;
        unsigned n_ext = 0;
        edges = 1;
        while (true) {
            float q[D];
            int key = 0;
            bool dead;
            if (phase == 0) {
                // virtual lane (lane, group) interleaved, so every pass of
                // every group spans the whole edge
                const int k = (lane * G0 + gw) * passes + p;
                const float kk = (float)((k < g ? k : g - 1) + 1);
#pragma unroll
                for (int c = 0; c < D; c++) q[c] = fmaf(kk, delta[c], nnv[c]);
                dead = k >= g;
            } else {
                // from chain pass wide_from on, the remaining poses spread
                // over pose groups: group gw takes poses base + 64 gw + lane
                shape(wide_ok && cpass >= d.wide_from ? groups_for(total - base) : 1);
                const int j = base + gw * kWave + lane;
                const int jj = j < total ? j : total - 1;
                const int i = jj / g, k = jj % g;
#pragma unroll
                for (int c = 0; c < D; c++) q[c] = fmaf((float)(k + 1), step[c], chain[i * D + c]);
                dead = j >= total;
                key = i;
            }
            ctx.coll = false;
            ctx.dead = dead;
            ctx.key = key;
            ctx.pass = npass;
            // poses of the team's pass (the same count on every wave)
            poses += phase == 0 ? min(kWave * G0, (g - 1 - p) / passes + 1) : min(total - base, kWave * groups);
            // the other parity's word is next read in the following pass,
            // after this pass's closing barrier
            if (tid == 0 && tm.waves > 1) stopw[(npass + 1) & 1] = 0x7fffffff;
            Robot::check(q, ctx);
            const uint64_t cm = ballot(ctx.coll);
            if (phase == 1) {
                // first colliding sub-edge: the smallest any wave found
                const int wfirst = cm ? __builtin_amdgcn_readlane(key, __builtin_ctzll(cm)) : 0x7fffffff;
                const int first = across_waves(wfirst, true, wbuf, npass++ & 1, tm);
                if (first != 0x7fffffff) {
                    n_ok = first;
                    break;
                }
                base += kWave * groups;
                cpass++;
                if (base >= total) { n_ok = n_chain; break; }
                continue;
            }
            if (across_waves(cm ? 1 : 0, false, wbuf, npass++ & 1, tm)) break;   // steered edge collides
            if (++p < passes) continue;
            // steered edge free: connect to the other tree (pRRTC.cu:478-518)
            flags |= kExtOk;
            float cfg[D];
#pragma unroll
            for (int c = 0; c < D; c++) cfg[c] = chain[c];
            const float *on = d.nodes + (size_t)o * D * d.cap;
            nn_o = nearest<D>(on, d.cap, o == 0 ? n0 : n1, cfg, bsq, red, tm);
            const float dist_o = sqrtf(bsq);
            n_ext = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)ceilf(dist_o / d.range));
            n_chain = __builtin_amdgcn_readfirstlane((int)min(n_ext, (unsigned)d.max_ext));
            if (tid < D) {
                const float vec = (on[(size_t)tid * d.cap + nn_o] - chain[tid]) / (float)n_ext;
                vecv[tid] = vec;
                step[tid] = vec / gf;
            }
            tm.sync();
            // chain configurations by sequential adds, thread i computes c_i
            for (int i0 = 1; i0 <= n_chain; i0 += NT) {
                const int i = i0 + tid;
                if (i <= n_chain) {
#pragma unroll
                    for (int c = 0; c < D; c++) {
                        const float vc = vecv[c];
                        float v = chain[c];
                        for (int r = 0; r < i; r++) v = v + vc;
                        chain[i * D + c] = v;
                    }
                }
            }
            tm.sync();
            total = n_chain * g;
            if (total == 0) { n_ok = 0; break; }
            phase = 1;
            base = 0;
        }
        if (flags & kExtOk) {
            edges += n_ok < n_chain ? n_ok + 1 : n_ok;
            if (n_ok == (int)n_ext) flags |= kConnected;
            // stage the new node and the extension nodes
            for (int i0 = 0; i0 <= n_ok; i0 += NT) {
                const int i = i0 + tid;
                if (i <= n_ok) {
#pragma unroll
                    for (int c = 0; c < D; c++) stage[(size_t)i * D + c] = chain[i * D + c];
                }
            }
        }
        tm.sync();
    }
    if (tid == 0) {
        meta[0] = t;
        meta[1] = flags;
        meta[2] = nn;
        meta[3] = n_ok;
        meta[4] = nn_o;
        __hip_atomic_fetch_add(&st->edge_checks, (unsigned long long)edges, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&st->pose_checks, (unsigned long long)poses, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Walk a parent chain from `cur` to its root; write rows (rs floats apart)
// when `out` is set.
template <int D>
__device__ int trace(const PlanDesc &d, int t, int cur, float *out, int rs, float *cost, int *root) {
#pragma clang fp contract(off)
    const float *tn = d.nodes + (size_t)t * D * d.cap;
    const int *par = d.parent + (size_t)t * d.cap;
    int n = 0;
    while (par[cur] != cur) {
        const int p = par[cur];
        if (out) {
            float acc = 0.0f;
            for (int k = 0; k < D; k++) {
                const float a = tn[(size_t)k * d.cap + cur];
                const float diff = a - tn[(size_t)k * d.cap + p];
                acc = fmaf(diff, diff, acc);
                out[(size_t)n * rs + k] = a;
            }
            *cost += sqrtf(acc);
        }
        n++;
        cur = p;
    }
    *root = cur;
    return n;
}

// The rows and segment lengths of the first (up to 64) rows of one walk,
// held one node per lane in `hop` (row j on lane j); `next` is the node after
// row 63 (a longer walk) and `root` the walk's root.  Rows go to `out`, and
// lane j's segment length |node_j - node_{j+1}| is returned.
template <int D>
__device__ float trace_rows(const PlanDesc &d, int t, int hop, int len, int next, int root, float *out, int rs) {
#pragma clang fp contract(off)
    const int lane = lane_id();
    const float *tn = d.nodes + (size_t)t * D * d.cap;
    const int up = __shfl_down(hop, 1);
    const int p = lane + 1 < len ? (lane + 1 < kWave ? up : next) : root;
    float acc = 0.0f;
    if (lane < len) {
        for (int k = 0; k < D; k++) {
            const float a = tn[(size_t)k * d.cap + hop];
            const float diff = a - tn[(size_t)k * d.cap + p];
            acc = fmaf(diff, diff, acc);
            out[(size_t)lane * rs + k] = a;
        }
    }
    return sqrtf(acc);
}

// Trace the connection of a round (pRRTC.cu:552-582) with the whole wave:
// lane 0 walks tree t from its new node and lane 1 the other tree from its
// nearest node, so every hop is one dependent load for both walks, and the
// first 64 nodes of each walk stay in registers (row j on lane j).  After the
// rows are reserved, those rows are written and their segments measured in
// parallel; the cost is summed in walk order (tree t's segments, then the
// other tree's), the order of a sequential trace, and a walk longer than 64
// rows continues on lane 0 from its 65th node.  Lane 0 holds the cost and
// writes the state.
template <int D>
__device__ void trace_connection(const PlanDesc &d, PlanState *st, PlanControl *ctrl, float *paths, int t,
                                 int cur_t, int cur_o) {
#pragma clang fp contract(off)
    const int lane = lane_id();
    const int *par = d.parent + (size_t)(lane == 0 ? t : 1 - t) * d.cap;
    int cur = lane == 0 ? cur_t : cur_o;
    bool walking = lane < 2;
    int len = 0;
    int hop_t = 0, hop_o = 0, next_t = -1, next_o = -1;
    while (true) {
        const int nxt = walking ? par[cur] : cur;
        const bool row = walking && nxt != cur;   // cur is not the root: a row of its walk
        const uint64_t m = ballot(row);
        if (m == 0) break;
        const int c0 = __builtin_amdgcn_readlane(cur, 0), c1 = __builtin_amdgcn_readlane(cur, 1);
        const int l0 = __builtin_amdgcn_readlane(len, 0), l1 = __builtin_amdgcn_readlane(len, 1);
        if (m & 1) {
            if (lane == l0) hop_t = c0;
            if (l0 == kWave) next_t = c0;
        }
        if (m & 2) {
            if (lane == l1) hop_o = c1;
            if (l1 == kWave) next_o = c1;
        }
        if (row) {
            len++;
            cur = nxt;
        } else {
            walking = false;
        }
    }
    const int tl = __builtin_amdgcn_readlane(len, 0), ol = __builtin_amdgcn_readlane(len, 1);
    const int root_t = __builtin_amdgcn_readlane(cur, 0), root_o = __builtin_amdgcn_readlane(cur, 1);
    int off = 0;
    if (lane == 0) off = atomic_add_agent(&ctrl->path_rows, tl + ol);
    off = __builtin_amdgcn_readfirstlane(off);
    if (off + tl + ol > ctrl->path_cap) {
        if (lane == 0) st->status = kStatusPathLimit;
        return;
    }
    const int rs = ctrl->path_stride > 0 ? ctrl->path_stride : D;
    float *out_t = paths + (size_t)off * rs, *out_o = paths + (size_t)(off + tl) * rs;
    const float seg_t = trace_rows<D>(d, t, hop_t, tl, next_t, root_t, out_t, rs);
    const float seg_o = trace_rows<D>(d, 1 - t, hop_o, ol, next_o, root_o, out_o, rs);
    float cost = 0.0f;
    int root;
    for (int j = 0; j < min(tl, kWave); j++) cost += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(seg_t), j));
    if (tl > kWave && lane == 0) trace<D>(d, t, next_t, out_t + (size_t)kWave * rs, rs, &cost, &root);
    for (int j = 0; j < min(ol, kWave); j++) cost += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(seg_o), j));
    if (ol > kWave && lane == 0) trace<D>(d, 1 - t, next_o, out_o + (size_t)kWave * rs, rs, &cost, &root);
    if (lane == 0) {
        st->t_tree = t;
        st->t_len = tl;
        st->o_len = ol;
        st->path_off = off;
        st->reached = t == 1 ? root_t : root_o;
        st->cost = cost;
        st->iters = st->round;
        st->status = kStatusSolved;
    }
}

__device__ __forceinline__ float dd_apply(const PlanDesc &d, float r, bool grow) {
#pragma clang fp contract(off)
    if (grow) return r == FLT_MAX ? r : r * (1 + d.dd_alpha);                         // pRRTC.cu:456-468
    return r == FLT_MAX ? d.dd_radius : fmaxf(r * (1.f - d.dd_alpha), d.dd_min_radius); // pRRTC.cu:584-598
}

// Commit a finished round (run by the wave that finished its last task).
// Returns the problem's new status.
template <class Robot>
__device__ int plan_commit(const PlanDesc &d, PlanState *st, int p, PlanControl *ctrl, float *paths) {
    constexpr int D = Robot::DOF;
    const int lane = lane_id();
    int sz[2] = {st->size[0], st->size[1]};
    int status = kStatusRunning;
    int win_slot = -1, win_last = -1, win_t = 0, win_nn_o = 0;
    for (int c0 = 0; c0 < d.nnc && status == kStatusRunning && win_slot < 0; c0 += kWave) {
        const int s = c0 + lane;
        const bool valid = s < d.nnc;
        int t = 0, flags = kSkip, nn = 0, n_ok = 0, nn_o = 0;
        if (valid) {
            const int *m = d.meta + (size_t)s * 8;
            t = m[0]; flags = m[1]; nn = m[2]; n_ok = m[3]; nn_o = m[4];
        }
        const bool skip = flags & kSkip;
        const bool ext_ok = flags & kExtOk;
        const bool conn = ext_ok && (flags & kConnected);
        const int count = ext_ok ? 1 + n_ok : 0;
        // positions: exclusive prefix over earlier slots of the same tree
        int off0 = (t == 0) ? count : 0, off1 = (t == 1) ? count : 0;
#pragma unroll
        for (int sh = 1; sh < kWave; sh <<= 1) {
            const int a0 = __shfl_up(off0, sh), a1 = __shfl_up(off1, sh);
            if (lane >= sh) { off0 += a0; off1 += a1; }
        }
        const int incl = t == 0 ? off0 : off1;
        const int start = (t == 0 ? sz[0] : sz[1]) + incl - count;
        const bool over = count > 0 && (start + count > d.max_samples || start + count > d.cap);
        const uint64_t cm = ballot(valid && conn);
        const uint64_t om = ballot(valid && over);
        const int first_conn = cm ? __builtin_ctzll(cm) : kWave;
        const int first_over = om ? __builtin_ctzll(om) : kWave;
        const int last_slot = min(first_conn, first_over);       // slots <= last_slot take effect
        // dynamic-domain updates in slot order (pRRTC.cu:456-468, 584-598):
        // equal (tree, node) keys are resolved in registers, lane by lane
        if (d.dd) {
            const bool has = valid && !skip && lane <= last_slot;
            const int key = (t << 28) | nn;
            float r = has ? d.radii[(size_t)t * d.cap + nn] : 0.0f;
            const uint64_t hm = ballot(has);
            for (uint64_t m = hm; m; m &= m - 1) {
                const int j = __builtin_ctzll(m);
                const int kj = __builtin_amdgcn_readlane(key, j);
                const int gj = __builtin_amdgcn_readlane((int)ext_ok, j);
                const float rj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(r), j));
                const float nv = dd_apply(d, rj, gj != 0);
                if (has && key == kj) r = nv;
            }
            if (has) d.radii[(size_t)t * d.cap + nn] = r;
        }
        if (first_over <= first_conn && first_over < kWave) {
            status = kStatusNodeLimit;
            break;
        }
        // insert nodes of slots <= last_slot
        if (valid && count > 0 && lane <= last_slot) {
            const float *stage = d.stage + (size_t)s * (d.max_ext + 1) * D;
            float *tn = d.nodes + (size_t)t * D * d.cap;
            int *par = d.parent + (size_t)t * d.cap;
            float *rad = d.radii + (size_t)t * d.cap;
            for (int i = 0; i < count; i++) {
                const int idx = start + i;
#pragma unroll
                for (int k = 0; k < D; k++) tn[(size_t)k * d.cap + idx] = stage[(size_t)i * D + k];
                par[idx] = i == 0 ? nn : idx - 1;
                rad[idx] = FLT_MAX;
            }
        }
        // sizes after this chunk
        int add0 = __builtin_amdgcn_readlane(off0, min(last_slot, kWave - 1));
        int add1 = __builtin_amdgcn_readlane(off1, min(last_slot, kWave - 1));
        sz[0] += add0;
        sz[1] += add1;
        if (first_conn < kWave) {
            win_slot = c0 + first_conn;
            win_t = __builtin_amdgcn_readlane(t, first_conn);
            win_nn_o = __builtin_amdgcn_readlane(nn_o, first_conn);
            const int ws = __builtin_amdgcn_readlane(start, first_conn);
            const int wc = __builtin_amdgcn_readlane(count, first_conn);
            win_last = ws + wc - 1;
        }
        __threadfence();   // this chunk's radii and nodes before the next chunk reads them
    }
    if (status == kStatusRunning && win_slot >= 0) {
        if (lane == 0) {
            st->size[0] = sz[0];
            st->size[1] = sz[1];
        }
        trace_connection<D>(d, st, ctrl, paths, win_t, win_last, win_nn_o);
        status = __builtin_amdgcn_readfirstlane(lane == 0 ? st->status : 0);
    } else if (lane == 0) {
        st->size[0] = sz[0];
        st->size[1] = sz[1];
        if (status == kStatusRunning && st->round >= d.max_iters) status = kStatusIterLimit;
        if (status == kStatusRunning && stop_requested(ctrl, p)) status = kStatusStopped;
        st->status = status;
    }
    status = __builtin_amdgcn_readfirstlane(status);
    __threadfence();
    if (lane == 0) {
        const unsigned long long key = ((unsigned long long)st->round << 32) | (unsigned)p;
        if (status == kStatusRunning) {
            const int round = st->round + 1;
            st->round = round;
            st->done_slots = 0;
            // a problem leaves the young class before its next round opens
            if (ctrl->aged_mask && round == ctrl->young_rounds)
                __hip_atomic_fetch_or(ctrl->aged_mask + (p >> 6), 1ull << (p & 63), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            store_release(&st->next_slot, 0);
            __hip_atomic_fetch_or(ctrl->open_mask + (p >> 6), 1ull << (p & 63), __ATOMIC_RELEASE,
                                  __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_max(&ctrl->hot_key, key + (1ull << 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            store_release(&st->next_slot, kClaimedOut);
            __hip_atomic_fetch_and(ctrl->run_mask + (p >> 6), ~(1ull << (p & 63)), __ATOMIC_RELEASE,
                                   __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_and(ctrl->open_mask + (p >> 6), ~(1ull << (p & 63)), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            // a finished problem hands the priority back to the next commit
            unsigned long long expect = key;
            __hip_atomic_compare_exchange_strong(&ctrl->hot_key, &expect, 0ull, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    return status;
}

// A batch whose problems have different robots runs one persistent grid:
// AnyRobot stands for all of them.  Its DOF (the largest) lays out the LDS
// task areas; each task, commit and problem setup dispatches on the
// problem's robot (a uniform branch), so a problem runs its own robot's code
// exactly as in a launch of that robot alone, and any free workgroup can
// take any problem's task.
constexpr int max_robot_dof() {
    int m = 0;
#define PRRTC_DOF_MAX(ID, R) m = R::DOF > m ? R::DOF : m;
    PRRTC_FOR_EACH_ROBOT(PRRTC_DOF_MAX)
#undef PRRTC_DOF_MAX
    return m;
}
struct AnyRobot {
    static constexpr int DOF = max_robot_dof();
};

// One task in the LDS task area `a` laid out for Robot::DOF (a robot's own
// layout fits in the larger AnyRobot one).
template <class Robot>
__device__ __forceinline__ void run_task(const PlanDesc &d, PlanState *st, int s, const TaskCtx &ctx, float *a,
                                         int lds_ext, const Team &tm) {
    constexpr int D = Robot::DOF;
    float *uv = a + (lds_ext + 1) * D, *red = a + (lds_ext + 5) * D;
    int *wbuf = reinterpret_cast<int *>(red + 2 * kMaxTaskWaves);
    if constexpr (std::is_same<Robot, AnyRobot>::value) {
        switch (d.robot) {
#define PRRTC_TASK_CASE(ID, R) case ID: plan_task<R>(d, st, s, ctx, a, uv, red, wbuf, tm); break;
            PRRTC_FOR_EACH_ROBOT(PRRTC_TASK_CASE)
#undef PRRTC_TASK_CASE
        }
    } else {
        plan_task<Robot>(d, st, s, ctx, a, uv, red, wbuf, tm);
    }
}

template <class Robot>
__device__ __forceinline__ void run_commit(const PlanDesc &d, PlanState *st, int p, PlanControl *ctrl, float *paths) {
    if constexpr (std::is_same<Robot, AnyRobot>::value) {
        switch (d.robot) {
#define PRRTC_COMMIT_CASE(ID, R) case ID: plan_commit<R>(d, st, p, ctrl, paths); break;
            PRRTC_FOR_EACH_ROBOT(PRRTC_COMMIT_CASE)
#undef PRRTC_COMMIT_CASE
        }
    } else {
        plan_commit<Robot>(d, st, p, ctrl, paths);
    }
}

template <class Robot>
__global__ __launch_bounds__(kMaxTaskWaves * kWave) __attribute__((amdgpu_waves_per_eu(PRRTC_PLAN_WAVES_PER_EU)))
void k_plan(const PlanDesc *__restrict__ descs, PlanState *states, PlanControl *ctrl, float *paths, int lds_scene_f4,
            int lds_ext) {
    extern __shared__ float4 lds[];
    // LDS: scene tables | gate exchange | one task area per wave (chain
    // rows, uniform vectors, cross-wave buffers) | the leader's decision.  A
    // task of the whole workgroup uses area 0; after a crowd claim wave w
    // runs its own task in area w.
    constexpr int D = Robot::DOF;
    const int W = blockDim.x / kWave;
    uint64_t *gbuf = reinterpret_cast<uint64_t *>(lds + lds_scene_f4);       // kMaxTaskWaves * kWave
    float *areas = reinterpret_cast<float *>(gbuf + kMaxTaskWaves * kWave);
    const int area = task_area_floats(D, lds_ext);
    int *ctl = reinterpret_cast<int *>(areas + (size_t)W * area);   // p, first slot, slots, priority, crowd
    TaskCtx ctx;
    ctx.independent = false;
    ctx.gbuf = gbuf;
    Obstacles scene_obs;   // the staged scene, all tables
    const bool leader = wave_id() == 0;   // wave 0 claims the tasks
    int staged = -1;
    int idle = 0, backoff = 1;
    const int P = ctrl->n_problems;
    const int prio_policy = ctrl->prio_policy;
    const int max_nnc = ctrl->max_nnc;
    const int crowd_above = W > 1 ? ctrl->crowd_above : 0;
    const unsigned long long slice = ctrl->slice_ticks;
    unsigned long long t0 = 0;
    if (slice && leader) {
        // one start time for the whole launch: the first workgroup sets it
        unsigned long long now = __builtin_amdgcn_s_memrealtime();
        if (lane_id() == 0) {
            unsigned long long expect = 0;
            __hip_atomic_compare_exchange_strong(&ctrl->t0, &expect, now, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            t0 = expect ? expect : now;
        }
        t0 = ((unsigned long long)__builtin_amdgcn_readfirstlane((int)(t0 >> 32)) << 32) |
             (unsigned)__builtin_amdgcn_readfirstlane((int)(t0 & 0xffffffffu));
    }
    const int words = ctrl->mask_words;
    int cursor = (int)(blockIdx.x % (unsigned)P);   // each workgroup starts at its own problem
    while (true) {
        if (leader) {
            int p_out = -1, s_out = -1, n_out = 1, hot = 0, crowd = 0;   // p_out < 0: the workgroup leaves
            int run_now = 0;   // running problems at the claim
            // Claim up to k consecutive slots of problem q if q runs and has
            // unclaimed slots in this round; returns the first slot (-1: none)
            // and their number in `got`.  A round the mask scan saw open is
            // claimed with the atomic alone: the counter is below nnc only
            // in an open round of a running problem (the commit resets it
            // for those only, after the round's state is written), and an
            // increment of a closed round's or a finished problem's counter
            // changes nothing (the next commit resets the former; the latter
            // holds 2^30, and only the short window between a last claim
            // or a final commit and its mask update sees a stale open bit).
            // A round not seen open is checked with plain loads first, so a
            // visit of a closed round costs no atomic.  The claim of the
            // last slot closes the problem in the open mask (before that
            // task can finish, so before the commit reopens it).
            auto try_claim = [&](int q, int k, bool seen_open, int &got) -> int {
                PlanState *st = states + q;
                const int nnc = descs[q].nnc;
                bool open = seen_open;
                if (!open) {
                    const bool live =
                        __builtin_amdgcn_readfirstlane(load_acquire(&st->status)) == kStatusRunning;
                    open = live && __builtin_amdgcn_readfirstlane(__hip_atomic_load(
                                       &st->next_slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) < nnc;
                }
                const int s = open ? wave_atomic_add(&st->next_slot, k) : nnc;
                if (s < 0 || s >= nnc) return -1;
                got = min(k, nnc - s);
                if (s + got == nnc && lane_id() == 0)
                    __hip_atomic_fetch_and(ctrl->open_mask + (q >> 6), ~(1ull << (q & 63)), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                return s;
            };
            while (true) {
                // running count and the next open problem from the cursor on
                // (young ones first on a server); closed and finished
                // problems are skipped through the masks, not visited
                // (and whether the furthest problem of a batch has an open
                // round: its hot_key is read first, its bit checked in the scan)
                const unsigned long long hk =
                    prio_policy == kPrioOldest
                        ? __hip_atomic_load(&ctrl->hot_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                        : 0ull;
                const int hp = __builtin_amdgcn_readfirstlane(hk ? (int)(hk & 0xffffffffu) : -1);
                const int ph = hp >= 0 && hp < P ? hp : -1;
                const MaskScan ms = scan_masks(ctrl, words, cursor, ph);
                run_now = ms.running;
                // the running problems can use at most running * nnc
                // workgroups: the rest leave, so a few long problems do not
                // share their counters with the whole grid
                if (ms.running == 0 || (long long)blockIdx.x >= (long long)ms.running * max_nnc) break;
                // a time-sliced launch ends between tasks; unfinished rounds
                // resume in the next launch from the state kept in device memory
                if (slice && __builtin_amdgcn_s_memrealtime() - t0 > slice) break;
                // many running problems (a crowd): each wave of the workgroup
                // takes its own slot and runs it alone, which spends no work
                // on splitting; few: the waves share one slot, which shortens
                // the round
                const int k = crowd_above > 0 && ms.running > crowd_above ? W : 1;
                int got = 1;
                // a batch serves the problem furthest along first, on all waves
                if (ms.want_open) {
                    const int s = try_claim(ph, 1, true, got);
                    if (s >= 0) {
                        p_out = ph;
                        s_out = s;
                        hot = 1;
                        idle = 0;
                        backoff = 1;
                        break;
                    }
                }
                const int q = __builtin_amdgcn_readfirstlane(ms.open_young >= 0 ? ms.open_young : ms.open);
                if (q >= 0) {
                    const int s = try_claim(q, k, true, got);
                    if (s >= 0) {
                        // the cursor stays: the next task is likely from the
                        // same problem, whose scene is already staged
                        cursor = q;
                        p_out = q;
                        s_out = s;
                        n_out = got;
                        crowd = k > 1;
                        idle = 0;
                        backoff = 1;
                        break;
                    }
                    // another workgroup took the last slot first: look on
                    cursor = q + 1 < P ? q + 1 : 0;
                    continue;
                }
                // every running round looks fully claimed: visit the running
                // problems in turn (which also finds a round whose open bit
                // was cleared late), and after a pass without work sleep
                // until a commit reopens one (exponential back-off, up to
                // 4 x 64 x 64 cycles)
                const int r = __builtin_amdgcn_readfirstlane(ms.first_run);
                const int s = r >= 0 ? try_claim(r, k, false, got) : -1;
                if (s >= 0) {
                    cursor = r;
                    p_out = r;
                    s_out = s;
                    n_out = got;
                    crowd = k > 1;
                    idle = 0;
                    backoff = 1;
                    break;
                }
                cursor = r + 1 < P ? r + 1 : 0;
                if (++idle >= ms.running) {
                    for (int i = 0; i < backoff; i++) __builtin_amdgcn_s_sleep(64);
                    backoff = backoff < 4 ? 2 * backoff : 4;
                    idle = 0;
                }
            }
            const int prio = p_out >= 0 ? task_priority(prio_policy, states[p_out].round, hot != 0) : 0;
            // a problem the host has stopped: its claimed tasks are skipped
            // (the round still completes and its commit ends the problem)
            const int halt = p_out >= 0 && stop_requested(ctrl, p_out) ? 1 : 0;
            if (lane_id() == 0) {
                ctl[0] = p_out;
                ctl[1] = s_out;
                ctl[2] = n_out;
                ctl[3] = prio;
                ctl[4] = crowd;
                ctl[5] = halt;
                ctl[6] = (long long)run_now * max_nnc <= ctrl->wide_slots;
            }
        }
        __syncthreads();
        const int p = ctl[0], s = ctl[1], n_slots = ctl[2];
        const bool crowd = ctl[4] != 0;
        if (p < 0) break;
        // issue priority of this task's waves on their SIMDs
        switch (ctl[3]) {
            case 3: __builtin_amdgcn_s_setprio(3); break;
            case 2: __builtin_amdgcn_s_setprio(2); break;
            case 1: __builtin_amdgcn_s_setprio(1); break;
            default: __builtin_amdgcn_s_setprio(0); break;
        }
        const PlanDesc &d = descs[p];
        PlanState *st = states + p;
        if (staged != p) {   // every wave is past its last use of the staged scene
            bind_scene<true>(d.scene, d.scene_f4, lds, scene_obs);
            staged = p;
        }
        // Without a crowd claim every wave runs the one slot and checks its
        // share of the obstacles; after one, wave w runs slot s + w on its
        // own with the whole scene (waves beyond the claimed slots idle).
        // One call site of plan_task and of plan_commit either way.
        const int w = wave_id();
        const bool runs = !crowd || w < n_slots;
        if (runs && ctl[5]) {
            // stopped by the host: the slot's outcome is a skip (no nodes, no
            // dynamic-domain update); one thread per slot writes it
            if (lane_id() == 0 && (crowd || w == 0)) {
                int *meta = d.meta + (size_t)(crowd ? s + w : s) * 8;
                meta[0] = 0;
                meta[1] = kSkip;
                meta[2] = meta[3] = meta[4] = 0;
            }
            __threadfence();
        } else if (runs) {
            ctx.obs = scene_obs;
            ctx.wide_ok = !crowd && ctl[6] != 0;
            float *a = areas + (crowd ? (size_t)w * area : 0);
            run_task<Robot>(d, st, crowd ? s + w : s, ctx, a, lds_ext, crowd ? team_wave() : team_workgroup());
            __threadfence();
        }
        // a shared task is written whole before its slot is counted
        if (!crowd) __syncthreads();
        // the wave that counts the round's last slot commits the round
        if (crowd ? runs : leader) {
            const int prev = wave_atomic_add(&st->done_slots, 1);
            if (prev == d.nnc - 1) run_commit<Robot>(d, st, p, ctrl, paths);
        }
        if (crowd) __syncthreads();   // ctl read by every wave before the leader claims again
    }
}

// Per-problem setup: tree roots, Halton states (pRRTC.cu:644-693).
template <class Robot>
__device__ __forceinline__ void init_problem(const PlanDesc &d, PlanState *st) {
    constexpr int D = Robot::DOF;
    const int lane = lane_id();
    for (int i = lane; i < 1 + d.n_goals; i += kWave) {
        const int t = i == 0 ? 0 : 1, idx = i == 0 ? 0 : i - 1;
        for (int k = 0; k < D; k++) d.nodes[(size_t)t * D * d.cap + (size_t)k * d.cap + idx] = d.init[(size_t)i * D + k];
        d.parent[(size_t)t * d.cap + idx] = idx;
        d.radii[(size_t)t * d.cap + idx] = FLT_MAX;
    }
    for (int s = lane; s < d.nnc; s += kWave) {
        float b[16];
        halton_bases(d.seed, s, D, b);
        float *hs = d.halton + (size_t)s * 3 * D;
        for (int k = 0; k < D; k++) { hs[k] = b[k]; hs[D + k] = 0.0f; hs[2 * D + k] = 1.0f; }
        d.tprev[s] = 0;
    }
    if (lane == 0) {
        *st = PlanState{};
        st->status = kStatusRunning;
        st->round = 1;
        st->size[0] = 1;
        st->size[1] = d.n_goals;
        st->next_slot = 0;
    }
}

template <class Robot>
__global__ __launch_bounds__(kWave) void k_plan_init(const PlanDesc *__restrict__ descs, PlanState *states,
                                                     const int *__restrict__ list) {
    const int e = list ? list[blockIdx.x] : (int)blockIdx.x;   // problem (table entry) to set up
    const PlanDesc &d = descs[e];
    PlanState *st = states + e;
    if constexpr (std::is_same<Robot, AnyRobot>::value) {
        switch (d.robot) {
#define PRRTC_SETUP_CASE(ID, R) case ID: init_problem<R>(d, st); break;
            PRRTC_FOR_EACH_ROBOT(PRRTC_SETUP_CASE)
#undef PRRTC_SETUP_CASE
        }
    } else {
        init_problem<Robot>(d, st);
    }
}

// The same rounds as plain grids, one kernel pair per round (the
// non-persistent form, replayed from a captured graph): k_round_tasks runs
// slot b % max_nnc of problem b / max_nnc on workgroup b, k_round_commit
// then commits every running problem's round.  Same tasks, same commit
// order, so the same results as k_plan.
template <class Robot>
__global__ __launch_bounds__(kMaxTaskWaves * kWave) __attribute__((amdgpu_waves_per_eu(PRRTC_PLAN_WAVES_PER_EU)))
void k_round_tasks(const PlanDesc *__restrict__ descs, PlanState *states, int max_nnc, int lds_scene_f4,
                   int lds_ext) {
    extern __shared__ float4 lds[];
    const int p = (int)(blockIdx.x / (unsigned)max_nnc), s = (int)(blockIdx.x % (unsigned)max_nnc);
    const PlanDesc &d = descs[p];
    PlanState *st = states + p;
    if (s >= d.nnc || st->status != kStatusRunning) return;   // the whole workgroup leaves
    constexpr int D = Robot::DOF;
    uint64_t *gbuf = reinterpret_cast<uint64_t *>(lds + lds_scene_f4);
    float *a = reinterpret_cast<float *>(gbuf + kMaxTaskWaves * kWave);   // task area 0 (k_plan's layout)
    TaskCtx ctx;
    ctx.independent = false;
    ctx.gbuf = gbuf;
    bind_scene<true>(d.scene, d.scene_f4, lds, ctx.obs);
    ctx.wide_ok = false;   // a grid of every running problem's slots: throughput-bound
    plan_task<Robot>(d, st, s, ctx, a, a + (lds_ext + 1) * D, a + (lds_ext + 5) * D,
                     reinterpret_cast<int *>(a + (lds_ext + 5) * D + 2 * kMaxTaskWaves), team_workgroup());
}

template <class Robot>
__global__ __launch_bounds__(kWave) void k_round_commit(const PlanDesc *__restrict__ descs, PlanState *states,
                                                        PlanControl *ctrl, float *paths) {
    const int p = blockIdx.x;
    PlanState *st = states + p;
    if (st->status != kStatusRunning) return;
    plan_commit<Robot>(descs[p], st, p, ctrl, paths);
}

// ------------------------------------------------------------------ host

struct Align {
    size_t off = 0;
    size_t take(size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    }
};

// Device memory, pinned staging, stream and timing events of one robot's
// group, reused across calls of this thread (grown on demand).
struct Workspace {
    char *dev = nullptr;
    size_t bytes = 0;
    char *host = nullptr;       // pinned staging for uploads / readback
    size_t host_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void ensure_stream() {
        if (stream) return;
        PRRTC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        PRRTC_HIP(hipEventCreate(&e0));
        PRRTC_HIP(hipEventCreate(&e1));
    }
    char *need_dev(size_t n) {
        if (n > bytes) {
            if (dev) (void)hipFree(dev);
            dev = nullptr;
            bytes = 0;
            PRRTC_HIP(hipMalloc(&dev, n));
            bytes = n;
        }
        return dev;
    }
    char *need_host(size_t n) {
        if (n > host_bytes) {
            if (host) (void)hipHostFree(host);
            host = nullptr;
            host_bytes = 0;
            PRRTC_HIP(hipHostMalloc(&host, n, hipHostMallocDefault));
            host_bytes = n;
        }
        return host;
    }
};

static Workspace &workspace(int robot) {
    static thread_local Workspace *ws[kRobotMixed + 1] = {};   // released with the process
    if (!ws[robot]) ws[robot] = new Workspace();
    ws[robot]->ensure_stream();
    return *ws[robot];
}

static int settings_error(const prrtc_settings *s) {
    if (!s) return 1;
    if (s->num_new_configs < 1 || s->num_new_configs > kMaxNnc) return 2;
    if (s->granularity < 1 || s->granularity > 4096) return 3;
    if (!(s->range > 0.0f) || !std::isfinite(s->range)) return 4;
    if (s->max_samples < 1 || s->max_iters < 1) return 5;
    if (s->balance < 0 || s->balance > 2) return 6;
    if (s->dynamic_domain && (!std::isfinite(s->dd_alpha) || !std::isfinite(s->dd_radius) ||
                              !std::isfinite(s->dd_min_radius)))
        return 7;
    return 0;
}

int robot_limits(int robot, float *lo, float *span) {
    switch (robot) {
#define PRRTC_LIM(ID, R) \
    case ID: for (int i = 0; i < R::DOF; i++) { lo[i] = R::lower(i); span[i] = R::span(i); } return R::DOF;
        PRRTC_FOR_EACH_ROBOT(PRRTC_LIM)
#undef PRRTC_LIM
        default: return -1;
    }
}

// Longest possible connect chain: the diagonal of the box spanned by the
// joint limits, the start and the goals, over `range`.
int max_extensions(const HostProblem &hp) {
    float lo[kMaxDof], span[kMaxDof];
    const int D = robot_limits(hp.robot, lo, span);
    double diag = 0.0;
    for (int k = 0; k < D; k++) {
        double a = lo[k], b = lo[k] + span[k];
        a = std::min(a, (double)hp.start[k]);
        b = std::max(b, (double)hp.start[k]);
        for (int i = 0; i < hp.n_goals; i++) {
            a = std::min(a, (double)hp.goals[(size_t)i * D + k]);
            b = std::max(b, (double)hp.goals[(size_t)i * D + k]);
        }
        diag += (b - a) * (b - a);
    }
    return (int)std::ceil(std::sqrt(diag) / hp.st->range) + 2;
}

BlockLayout block_layout(const HostProblem &hp) {
    const int D = robot_dof(hp.robot);
    const int nnc = hp.st->num_new_configs;
    BlockLayout L;
    L.scene_floats = (int)scene_packed_floats(hp.scene);
    // storage rounded up to whole 1024-node scan chunks (16-byte aligned SoA
    // columns, the nearest-node scan reads whole chunks); the node limit
    // stays max_samples
    L.cap = (std::max(hp.st->max_samples, hp.n_goals) + kScanChunk - 1) & ~(kScanChunk - 1);
    L.max_ext = max_extensions(hp);
    Align a;
    L.scene = a.take(0);
    a.off += sizeof(float) * (size_t)L.scene_floats;   // init rows follow the scene directly
    L.init = a.off;
    a.off += sizeof(float) * (size_t)(1 + hp.n_goals) * D;
    L.upload_bytes = a.off;
    a.off = (a.off + 255) & ~size_t(255);
    L.nodes = a.take(sizeof(float) * 2 * (size_t)D * L.cap);
    L.parent = a.take(sizeof(int) * 2 * (size_t)L.cap);
    L.radii = a.take(sizeof(float) * 2 * (size_t)L.cap);
    L.halton = a.take(sizeof(float) * 3 * (size_t)D * nnc);
    L.tprev = a.take(sizeof(int) * (size_t)nnc);
    L.meta = a.take(sizeof(int) * 8 * (size_t)nnc);
    L.stage = a.take(sizeof(float) * (size_t)nnc * (L.max_ext + 1) * D);
    L.bytes = a.off;
    return L;
}

void block_upload_image(const HostProblem &hp, const BlockLayout &L, char *host) {
    const int D = robot_dof(hp.robot);
    scene_pack(hp.scene, reinterpret_cast<float *>(host + L.scene));
    float *ini = reinterpret_cast<float *>(host + L.init);
    std::memcpy(ini, hp.start, sizeof(float) * D);
    std::memcpy(ini + D, hp.goals, sizeof(float) * (size_t)hp.n_goals * D);
}

// Chain pass from which a task's waves may take distinct poses (negative:
// never; PRRTC_WIDE_FROM overrides the robot's default).  Whether they do is
// decided per task (PlanControl::wide_slots).  A wide pass saves the
// forward kinematics the waves of a shared pass compute alike and the gate
// exchange; Fetch's checks are dominated by its bounded leaf groups rather
// than its kinematics, and its solves measured slower with wide passes (48
// C2 problems alone at nnc 128: 748 -> 923 ms of kernel; the mix's Fetch
// problems 197 -> 214 ms), Panda's and Baxter's faster (C0 16.5 -> 14.6 ms,
// the C1 plans 1582 -> 1461 ms, the mix's Panda 2236 -> 2110 and Baxter
// 1016 -> 965 ms).
static int wide_from(int robot) {
    static const int env = [] {
        const char *e = std::getenv("PRRTC_WIDE_FROM");
        return e ? std::atoi(e) : INT32_MIN;
    }();
    if (env != INT32_MIN) return env;
    return robot == PRRTC_ROBOT_FETCH ? -1 : 0;
}

// Chain passes go wide only while the running problems' rounds leave half
// the CUs free: such rounds wait for their slowest task.  A round that fills
// the machine is bound by throughput, and the poses a wide pass adds (past a
// chain's first collision, on waves without poses) cost more than its
// shorter passes save (the 256-slot C2 plain solves 15% slower, the C2
// batch's tail 13%).
int plan_wide_slots() { return num_cus() / 2; }

// Whether a steered edge of more than 64 poses spreads them over pose groups
// (PRRTC_WIDE_EDGE overrides).
static int wide_edge() {
    static const int v = [] {
        const char *e = std::getenv("PRRTC_WIDE_EDGE");
        return e ? std::atoi(e) : 1;
    }();
    return v;
}

void fill_desc(PlanDesc &d, const HostProblem &hp, const BlockLayout &L, char *dev) {
    d = PlanDesc{};
    d.wide_from = wide_from(hp.robot);
    d.wide_edge = wide_edge();
    d.nnc = hp.st->num_new_configs;
    d.g = hp.st->granularity;
    d.max_iters = hp.st->max_iters;
    d.max_samples = hp.st->max_samples;
    d.balance = hp.st->balance;
    d.dd = hp.st->dynamic_domain ? 1 : 0;
    d.n_goals = hp.n_goals;
    d.robot = hp.robot;
    d.cap = L.cap;
    d.max_ext = L.max_ext;
    d.scene_f4 = (L.scene_floats - 8) / 4;
    d.range = hp.st->range;
    d.tree_ratio = hp.st->tree_ratio;
    d.dd_alpha = hp.st->dd_alpha;
    d.dd_radius = hp.st->dd_radius;
    d.dd_min_radius = hp.st->dd_min_radius;
    d.scene = reinterpret_cast<const float *>(dev + L.scene);
    d.nodes = reinterpret_cast<float *>(dev + L.nodes);
    d.parent = reinterpret_cast<int *>(dev + L.parent);
    d.radii = reinterpret_cast<float *>(dev + L.radii);
    d.halton = reinterpret_cast<float *>(dev + L.halton);
    d.tprev = reinterpret_cast<int *>(dev + L.tprev);
    d.meta = reinterpret_cast<int *>(dev + L.meta);
    d.stage = reinterpret_cast<float *>(dev + L.stage);
    d.init = reinterpret_cast<const float *>(dev + L.init);
    d.seed = hp.st->seed;
}

int assemble_path(const HostProblem &hp, const PlanState &ps, const float *rows, int rs, int path_cap, float *out) {
    const int D = robot_dof(hp.robot);
    const int n_rows = ps.t_len + ps.o_len + 2;
    if (n_rows > path_cap) return -1;
    const float *tr = rows + (size_t)ps.path_off * rs;      // t-trace then o-trace
    const float *p1 = ps.t_tree == 1 ? tr : tr + (size_t)ps.t_len * rs;
    const float *p0 = ps.t_tree == 1 ? tr + (size_t)ps.t_len * rs : tr;
    const int n1 = ps.t_tree == 1 ? ps.t_len : ps.o_len;
    const int n0 = ps.t_tree == 1 ? ps.o_len : ps.t_len;
    int row = 0;
    std::memcpy(out + (size_t)row++ * D, hp.goals + (size_t)ps.reached * D, sizeof(float) * D);
    for (int k = n1 - 1; k >= 0; k--) std::memcpy(out + (size_t)row++ * D, p1 + (size_t)k * rs, sizeof(float) * D);
    for (int k = 0; k < n0; k++) std::memcpy(out + (size_t)row++ * D, p0 + (size_t)k * rs, sizeof(float) * D);
    std::memcpy(out + (size_t)row++ * D, hp.start, sizeof(float) * D);
    return row;
}

int plan_dof(int robot) { return robot == kRobotMixed ? AnyRobot::DOF : robot_dof(robot); }

void launch_plan_init(int robot, const PlanDesc *d_desc, PlanState *d_state, const int *d_list, int n,
                      hipStream_t s) {
    if (n <= 0) return;
    switch (robot) {
#define PRRTC_INIT_CASE(ID, R) \
    case ID: hipLaunchKernelGGL(k_plan_init<R>, dim3(n), dim3(kWave), 0, s, d_desc, d_state, d_list); break;
        PRRTC_FOR_EACH_ROBOT(PRRTC_INIT_CASE)
        PRRTC_INIT_CASE(kRobotMixed, AnyRobot)
#undef PRRTC_INIT_CASE
    }
    PRRTC_HIP(hipGetLastError());
}

// Dynamic LDS of a planner grid: scene, gate exchange, one task area per
// wave, the leader's decision (the layout k_plan and k_round_tasks use).
static size_t plan_lds_bytes(int D, int lds_scene_f4, int max_ext, int waves) {
    return (size_t)lds_scene_f4 * sizeof(float4) + kMaxTaskWaves * kWave * sizeof(uint64_t) +
           ((size_t)waves * task_area_floats(D, max_ext) + kCtlInts) * sizeof(float);
}

void launch_plan_slice(int robot, const PlanDesc *d_desc, PlanState *d_state, PlanControl *d_ctrl, float *d_paths,
                       int grid, int waves, int lds_scene_f4, int max_ext, hipStream_t s) {
    const size_t lds = plan_lds_bytes(plan_dof(robot), lds_scene_f4, max_ext, waves);
    switch (robot) {
#define PRRTC_SLICE_CASE(ID, R)                                                                              \
    case ID:                                                                                                 \
        hipLaunchKernelGGL(k_plan<R>, dim3(grid), dim3(waves * kWave), lds, s, d_desc, d_state, d_ctrl,      \
                           d_paths, lds_scene_f4, max_ext);                                                  \
        break;
        PRRTC_FOR_EACH_ROBOT(PRRTC_SLICE_CASE)
        PRRTC_SLICE_CASE(kRobotMixed, AnyRobot)
#undef PRRTC_SLICE_CASE
    }
    PRRTC_HIP(hipGetLastError());
}

// Run rounds until no problem runs: `per_graph` rounds (a task grid and a
// commit grid each) and a copy of the running mask to the host are
// captured once as a graph and replayed; the host reads the mask after each
// replay.  At most max_rounds rounds (every running problem advances one
// per round and stops at its max_iters).
void run_round_graph(int robot, const PlanDesc *d_desc, PlanState *d_state, PlanControl *d_ctrl, float *d_paths,
                     int P, int max_nnc, int waves, int lds_scene_f4, int max_ext, const unsigned long long *d_run,
                     unsigned long long *h_run, int words, long long max_rounds, hipStream_t s) {
    constexpr int per_graph = 8;
    const size_t lds = plan_lds_bytes(robot_dof(robot), lds_scene_f4, max_ext, 1);   // task area 0 only
    const long long tasks = (long long)P * max_nnc;
    if (tasks > INT32_MAX) throw HipError{hipErrorInvalidValue, "run_round_graph: 2^31 tasks per round"};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    PRRTC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int r = 0; r < per_graph; r++) {
        switch (robot) {
#define PRRTC_ROUND_CASE(ID, R)                                                                                  \
    case ID:                                                                                                     \
        hipLaunchKernelGGL(k_round_tasks<R>, dim3((unsigned)tasks), dim3(waves * kWave), lds, s, d_desc, d_state, \
                           max_nnc, lds_scene_f4, max_ext);                                                      \
        hipLaunchKernelGGL(k_round_commit<R>, dim3(P), dim3(kWave), 0, s, d_desc, d_state, d_ctrl, d_paths);     \
        break;
            PRRTC_FOR_EACH_ROBOT(PRRTC_ROUND_CASE)
#undef PRRTC_ROUND_CASE
        }
    }
    (void)hipMemcpyAsync(h_run, d_run, sizeof(unsigned long long) * words, hipMemcpyDeviceToHost, s);
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    PRRTC_HIP(ec);
    struct Release {
        hipGraph_t &g;
        hipGraphExec_t &e;
        ~Release() {
            if (e) (void)hipGraphExecDestroy(e);
            if (g) (void)hipGraphDestroy(g);
        }
    } release{graph, exec};
    PRRTC_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    for (long long done = 0; done < max_rounds; done += per_graph) {
        PRRTC_HIP(hipGraphLaunch(exec, s));
        PRRTC_HIP(hipStreamSynchronize(s));
        bool any = false;
        for (int w = 0; w < words && !any; w++) any = h_run[w] != 0ull;
        if (!any) break;
    }
}

// Launch shape of each robot's persistent grids (prrtc_set_tuning; 0 = default).
static std::atomic<int> g_waves_per_cu[kNumRobots] = {};
static std::atomic<int> g_task_waves[kNumRobots] = {};
static std::atomic<int> g_batch_waves[kNumRobots] = {};
static std::atomic<int> g_crowd_above[kNumRobots] = {};

static int env_int(const char *name, int lo, int hi) {
    const char *e = std::getenv(name);
    const int n = e ? std::atoi(e) : 0;
    return n >= lo && n <= hi ? n : 0;
}

// waves of a task: 1, 2 or 4 (a task's obstacle shares are cut by shifts)
static int pow2_waves(int w) { return w == 3 ? 0 : w; }

static int tuned(const std::atomic<int> *table, int robot, int env, int dflt) {
    const int v = robot >= 0 && robot < kNumRobots ? table[robot].load(std::memory_order_relaxed) : 0;
    return v > 0 ? v : (env > 0 ? env : dflt);
}

int plan_task_waves(int robot) {
    static const int env = pow2_waves(env_int("PRRTC_TASK_WAVES", 1, kMaxTaskWaves));
    return tuned(g_task_waves, robot, env, kDefaultTaskWaves);
}

int plan_batch_waves(int robot) {
    static const int env = pow2_waves(env_int("PRRTC_BATCH_WAVES", 1, kMaxTaskWaves));
    return tuned(g_batch_waves, robot, env, kDefaultBatchWaves);
}

int plan_waves_per_cu(int robot) {
    static const int env = env_int("PRRTC_WAVES_PER_CU", 1, kMaxWavesPerCu);
    return tuned(g_waves_per_cu, robot, env, kDefaultWavesPerCu);
}

int plan_tail_problems(int robot, int max_nnc) {
    // a grid of one-wave tasks holds every slot of this many problems at once
    return std::max(1, num_cus() * plan_waves_per_cu(robot) / std::max(1, max_nnc));
}

// The crowd threshold as set (prrtc_set_tuning, else PRRTC_CROWD_ABOVE):
// > 0 a count, < 0 off, 0 the default.
static int crowd_setting(int robot) {
    static const int env = [] {
        const char *e = std::getenv("PRRTC_CROWD_ABOVE");
        return e ? std::atoi(e) : 0;
    }();
    const int v = robot >= 0 && robot < kNumRobots ? g_crowd_above[robot].load(std::memory_order_relaxed) : 0;
    return v != 0 ? v : env;
}

int plan_crowd_above(int robot, int max_nnc) {
    // by default a crowd is more problems than the resident planner waves
    // (kResidentPlanWavesPerCu per CU) can serve with one slot each
    const int v = crowd_setting(robot);
    if (v < 0) return 0;
    return v > 0 ? v : std::max(1, num_cus() * kResidentPlanWavesPerCu / std::max(1, max_nnc));
}

int plan_grid(int robot, int waves, long long slots) {
    const long long per_cu = std::max(1, plan_waves_per_cu(robot) / waves);
    return (int)std::max<long long>(1, std::min<long long>(slots, (long long)num_cus() * per_cu));
}

// Path rows read back together with a launch's states.
constexpr int kHeadRows = 512;

// Solve problems that share one robot in one persistent launch.
// A robot group in flight: its problems and where its state lives.
struct GroupRun {
    std::vector<HostProblem> probs;
    int robot = -1;              // the problems' robot, or kRobotMixed
    bool graph_rounds = false;   // graph-replayed round grids instead of one persistent grid
    Workspace *ws = nullptr;
    char *dev = nullptr, *host = nullptr;
    size_t o_state = 0, o_ctrl = 0, o_paths = 0;
    int head_rows = 0;           // path rows read back with the states (the host image holds them)
};

// Upload a group, launch its persistent grid and queue the state readback
// on the group's own stream (a mixed batch is one group; per-robot groups,
// where used, overlap on the device).
static void start_group(GroupRun &gr, int path_cap) {
    const std::vector<HostProblem> &probs = gr.probs;
    const int P = (int)probs.size();
    const int robot = gr.robot;
    const int D = plan_dof(robot);
    Workspace &ws = workspace(robot);
    // a mixed launch takes, of each launch shape, the largest value among
    // its problems' robots
    std::vector<int> robots;
    for (const HostProblem &hp : probs)
        if (std::find(robots.begin(), robots.end(), hp.robot) == robots.end()) robots.push_back(hp.robot);
    auto largest = [&](auto f) {
        int v = 0;
        for (int r : robots) v = std::max(v, f(r));
        return v;
    };
    gr.ws = &ws;
    // --- layout: what is uploaded (descs, masks, control), then what is
    // read back (control, states, the path rows), then the problems' blocks;
    // the host image mirrors the first two parts, so each is one copy
    std::vector<BlockLayout> lay(P);
    int64_t path_rows = 0;
    for (int i = 0; i < P; i++) {
        lay[i] = block_layout(probs[i]);
        path_rows += std::max(0, path_cap - 2);
    }
    Align dl;
    const size_t o_desc = dl.take(sizeof(PlanDesc) * P);
    const int words = (P + 63) / 64;
    const size_t o_mask = dl.take(sizeof(unsigned long long) * 2 * words);   // running | open
    const size_t o_ctrl = dl.take(sizeof(PlanControl));
    const size_t o_state = dl.take(sizeof(PlanState) * P);
    const size_t o_paths = dl.take(sizeof(float) * (size_t)path_rows * D);
    std::vector<size_t> o_block(P);
    int lds_f4 = 0, ext_all = 0;
    for (int i = 0; i < P; i++) {
        lds_f4 = std::max(lds_f4, (lay[i].scene_floats - 8) / 4);
        ext_all = std::max(ext_all, lay[i].max_ext);
        o_block[i] = dl.take(lay[i].bytes);
    }
    char *dev = ws.need_dev(dl.off);
    // the first kHeadRows path rows come back with the states: a solve whose
    // path fits needs no second copy (and no second wait)
    const int head_rows = (int)std::min<int64_t>(path_rows, kHeadRows);
    const size_t up_end = o_ctrl + sizeof(PlanControl);
    const size_t read_end = o_paths + sizeof(float) * (size_t)head_rows * D;
    Align hl;
    hl.take(read_end);   // the host image of [0, read_end)
    std::vector<size_t> h_block(P);
    for (int i = 0; i < P; i++) h_block[i] = hl.take(lay[i].upload_bytes);
    const size_t h_run = hl.take(sizeof(unsigned long long) * words);
    char *host = ws.need_host(hl.off);
    PlanDesc *hd = reinterpret_cast<PlanDesc *>(host + o_desc);
    for (int i = 0; i < P; i++) {
        fill_desc(hd[i], probs[i], lay[i], dev + o_block[i]);
        block_upload_image(probs[i], lay[i], host + h_block[i]);
    }
    PlanControl *hc = reinterpret_cast<PlanControl *>(host + o_ctrl);
    *hc = PlanControl{};
    hc->n_problems = P;
    hc->mask_words = words;
    hc->run_mask = reinterpret_cast<unsigned long long *>(dev + o_mask);
    hc->open_mask = hc->run_mask + words;   // every problem starts with round 0 open
    unsigned long long *hm = reinterpret_cast<unsigned long long *>(host + o_mask);
    for (int w = 0; w < words; w++) hm[w] = hm[words + w] = (w + 1) * 64 <= P ? ~0ull : (1ull << (P - w * 64)) - 1;
    hc->path_cap = (int)std::min<int64_t>(path_rows, INT32_MAX);
    hc->path_stride = D;
    for (int i = 0; i < P; i++) hc->max_nnc = std::max(hc->max_nnc, probs[i].st->num_new_configs);
    hc->prio_policy = kPrioOldest;
    hc->wide_slots = plan_wide_slots();
    hc->crowd_above = P > 1 ? largest([&](int r) { return plan_crowd_above(r, hc->max_nnc); }) : 0;
    // --- upload (descs, masks and control in one copy; scenes and init rows
    // per problem): the control block is complete before the copy is queued
    // (the copy reads the pinned staging asynchronously)
    const hipStream_t s = ws.stream;
    PRRTC_HIP(hipMemcpyAsync(dev + o_desc, host + o_desc, up_end - o_desc, hipMemcpyHostToDevice, s));
    for (int i = 0; i < P; i++)
        PRRTC_HIP(hipMemcpyAsync(dev + o_block[i], host + h_block[i], lay[i].upload_bytes, hipMemcpyHostToDevice, s));
    // --- launch shape: a single problem runs its tasks on waves_per_task
    // waves (latency), a batch on batch_waves_per_task (throughput); one
    // persistent grid either way
    long long total_slots = 0;
    for (int i = 0; i < P; i++) total_slots += probs[i].st->num_new_configs;
    const int waves = P == 1 ? plan_task_waves(robot) : largest(plan_batch_waves);
    const PlanDesc *d_desc = reinterpret_cast<const PlanDesc *>(dev + o_desc);
    PlanState *d_state = reinterpret_cast<PlanState *>(dev + o_state);
    PlanControl *d_ctrl = reinterpret_cast<PlanControl *>(dev + o_ctrl);
    float *d_paths = reinterpret_cast<float *>(dev + o_paths);
    PRRTC_HIP(hipEventRecord(ws.e0, s));
    launch_plan_init(robot, d_desc, d_state, nullptr, P, s);
    if (gr.graph_rounds) {
        long long max_rounds = 0;
        for (int i = 0; i < P; i++) max_rounds = std::max<long long>(max_rounds, probs[i].st->max_iters + 1);
        run_round_graph(robot, d_desc, d_state, d_ctrl, d_paths, P, hc->max_nnc, waves, lds_f4, ext_all,
                        hc->run_mask, reinterpret_cast<unsigned long long *>(host + h_run), words, max_rounds, s);
    } else {
        const int grid = largest([&](int r) { return plan_grid(r, waves, total_slots); });
        launch_plan_slice(robot, d_desc, d_state, d_ctrl, d_paths, grid, waves,
                          lds_f4, ext_all, s);
    }
    PRRTC_HIP(hipEventRecord(ws.e1, s));
    // --- queue the readback of control, states and the first path rows
    PRRTC_HIP(hipMemcpyAsync(host + o_ctrl, dev + o_ctrl, read_end - o_ctrl, hipMemcpyDeviceToHost, s));
    gr.dev = dev;
    gr.host = host;
    gr.o_state = o_state;
    gr.o_ctrl = o_ctrl;
    gr.o_paths = o_paths;
    gr.head_rows = head_rows;
}

// Wait for a started group, read the used path rows, write its results.
static void finish_group(GroupRun &gr, prrtc_result *results, float *paths, int path_cap, int path_stride,
                         int64_t *kernel_ns, int *rounds) {
    const std::vector<HostProblem> &probs = gr.probs;
    const int P = (int)probs.size();
    const int D = plan_dof(gr.robot);   // floats per path row (the launch's path_stride)
    Workspace &ws = *gr.ws;
    const hipStream_t s = ws.stream;
    PRRTC_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    PRRTC_HIP(hipEventElapsedTime(&ms, ws.e0, ws.e1));
    *kernel_ns = std::max(*kernel_ns, (int64_t)(ms * 1e6));   // groups overlap: the longest one
    const PlanState *hs = reinterpret_cast<const PlanState *>(gr.host + gr.o_state);
    const PlanControl *hctl = reinterpret_cast<const PlanControl *>(gr.host + gr.o_ctrl);
    const int used = std::min(hctl->path_rows, hctl->path_cap);
    // rows beyond those read back with the states need a copy of their own
    std::vector<float> more;
    const float *rows = reinterpret_cast<const float *>(gr.host + gr.o_paths);
    if (used > gr.head_rows) {
        more.resize((size_t)used * D);
        PRRTC_HIP(hipMemcpyAsync(more.data(), gr.dev + gr.o_paths, sizeof(float) * more.size(),
                                 hipMemcpyDeviceToHost, s));
        PRRTC_HIP(hipStreamSynchronize(s));
        rows = more.data();
    }
    // --- results, path goal -> start (pRRTC.cu:715-746)
    for (int i = 0; i < P; i++) {
        const HostProblem &hp = probs[i];
        const PlanState &ps = hs[i];
        prrtc_result &r = results[hp.index];
        r.start_tree_size = ps.size[0];
        r.goal_tree_size = ps.size[1];
        r.edge_checks = (int64_t)ps.edge_checks;
        r.pose_checks = (int64_t)ps.pose_checks;
        *rounds = std::max(*rounds, ps.round);
        r.solved = ps.status == kStatusSolved;
        r.status = r.solved ? PRRTC_OK : (ps.status == kStatusPathLimit ? PRRTC_EOVERFLOW : PRRTC_OK);
        r.iters = r.solved ? ps.iters : 0;
        r.cost = r.solved ? ps.cost : 0.0f;
        r.path_nodes = 0;
        r.path_length = 0;
        if (!r.solved) continue;
        const int row = assemble_path(hp, ps, rows, D, path_cap, paths + (size_t)hp.index * path_stride);
        if (row < 0) {
            r.solved = 0;
            r.status = PRRTC_EOVERFLOW;
            continue;
        }
        r.path_nodes = row;
        r.path_length = ps.t_len + ps.o_len;
    }
}

int check_problem(const HostProblem &hp, const char *fn) {
    const int D = robot_dof(hp.robot);
    if (D < 0) { set_error("%s: unknown robot id %d", fn, hp.robot); return PRRTC_EINVAL; }
    if (!hp.start || !hp.goals || hp.n_goals < 1) { set_error("%s: start/goals missing", fn); return PRRTC_EINVAL; }
    for (int k = 0; k < D; k++)
        if (!std::isfinite(hp.start[k])) { set_error("%s: non-finite start", fn); return PRRTC_EINVAL; }
    for (int k = 0; k < D * hp.n_goals; k++)
        if (!std::isfinite(hp.goals[k])) { set_error("%s: non-finite goal", fn); return PRRTC_EINVAL; }
    if (!scene_valid(hp.scene)) { set_error("%s: invalid scene", fn); return PRRTC_EINVAL; }
    if (scene_packed_floats(hp.scene) - 8 > kMaxLdsScene / sizeof(float)) {
        set_error("%s: scene tables exceed %zu bytes", fn, kMaxLdsScene);
        return PRRTC_EINVAL;
    }
    if (int e = settings_error(hp.st)) { set_error("%s: invalid settings (code %d)", fn, e); return PRRTC_EINVAL; }
    if (hp.n_goals > hp.st->max_samples) { set_error("%s: more goals than max_samples", fn); return PRRTC_EINVAL; }
    if (max_extensions(hp) > kMaxExt) {
        set_error("%s: range %g too small for this robot (connect chains over %d steps)", fn, hp.st->range, kMaxExt);
        return PRRTC_EINVAL;
    }
    return PRRTC_OK;
}

static bool mixed_grid() {
    static const bool on = [] {
        const char *e = std::getenv("PRRTC_MIXED_GRID");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

static int solve_many(std::vector<HostProblem> &probs, prrtc_result *results, float *paths, int path_cap,
                      int path_stride, prrtc_batch_stats *stats, const char *fn, bool graph_rounds = false) {
    const int64_t t0 = now_ns();
    for (auto &hp : probs) {
        int rc = check_problem(hp, fn);
        if (rc) return rc;
    }
    if (path_cap < 2 || !results || !paths) {
        set_error("%s: result/path buffers missing or path_cap < 2", fn);
        return PRRTC_EINVAL;
    }
    std::map<int, std::vector<HostProblem>> groups;
    for (auto &hp : probs) groups[hp.robot].push_back(hp);
    // Problems of several robots share one persistent grid (kRobotMixed):
    // any workgroup takes any problem's task, so one robot's long tail runs
    // beside another's bulk.  The graph-replayed comparator, and
    // PRRTC_MIXED_GRID=0, keep one grid per robot, each on its own stream.
    if (groups.size() > 1 && !graph_rounds && mixed_grid()) {
        groups.clear();
        groups[kRobotMixed] = probs;
    }
    int64_t kernel_ns = 0;
    int rounds = 0;
    std::vector<GroupRun> runs(groups.size());
    size_t k = 0;
    for (auto &kv : groups) {
        runs[k].probs = std::move(kv.second);
        runs[k].robot = kv.first;
        runs[k].graph_rounds = graph_rounds;
        start_group(runs[k++], path_cap);
    }
    for (GroupRun &gr : runs) finish_group(gr, results, paths, path_cap, path_stride, &kernel_ns, &rounds);
    const int64_t wall = now_ns() - t0;
    int solved = 0;
    int64_t edges = 0, poses = 0;
    for (auto &hp : probs) {
        results[hp.index].wall_ns = wall;
        results[hp.index].kernel_ns = kernel_ns;
        results[hp.index].copy_ns = 0;
        solved += results[hp.index].solved;
        edges += results[hp.index].edge_checks;
        poses += results[hp.index].pose_checks;
    }
    if (stats) {
        stats->wall_ns = wall;
        stats->kernel_ns = kernel_ns;
        stats->edge_checks = edges;
        stats->pose_checks = poses;
        stats->solved = solved;
        stats->rounds = rounds;
    }
    return PRRTC_OK;
}

}  // namespace prrtc

using namespace prrtc;

extern "C" {

int prrtc_solve(int robot, const float *start, const float *goals, int n_goals, const prrtc_scene *scene,
                const prrtc_settings *settings, prrtc_result *result, float *path, int path_cap) {
    return guarded("prrtc_solve", [&]() -> int {
        if (!result) { set_error("prrtc_solve: result is NULL"); return PRRTC_EINVAL; }
        *result = prrtc_result{};
        std::vector<HostProblem> probs{HostProblem{robot, start, goals, n_goals, scene, settings, 0}};
        return solve_many(probs, result, path, path_cap, 0, nullptr, "prrtc_solve");
    });
}

int prrtc_set_tuning(int robot, const prrtc_tuning *t) {
    auto bad = [](int v, int hi) { return v < 0 || v > hi; };
    if (!t || robot < -1 || robot >= kNumRobots || bad(t->waves_per_cu, kMaxWavesPerCu) ||
        bad(t->waves_per_task, kMaxTaskWaves) || bad(t->batch_waves_per_task, kMaxTaskWaves) ||
        t->waves_per_task == 3 || t->batch_waves_per_task == 3) {   // a task's waves: 1, 2 or 4
        set_error("prrtc_set_tuning: invalid argument");
        return PRRTC_EINVAL;
    }
    for (int r = 0; r < kNumRobots; r++) {
        if (robot >= 0 && r != robot) continue;
        g_waves_per_cu[r].store(t->waves_per_cu, std::memory_order_relaxed);
        g_task_waves[r].store(t->waves_per_task, std::memory_order_relaxed);
        g_batch_waves[r].store(t->batch_waves_per_task, std::memory_order_relaxed);
        g_crowd_above[r].store(t->crowd_above, std::memory_order_relaxed);
    }
    return PRRTC_OK;
}

int prrtc_get_tuning(int robot, prrtc_tuning *t) {
    if (!t || robot < 0 || robot >= kNumRobots) {
        set_error("prrtc_get_tuning: invalid argument");
        return PRRTC_EINVAL;
    }
    t->waves_per_cu = plan_waves_per_cu(robot);
    t->waves_per_task = plan_task_waves(robot);
    t->batch_waves_per_task = plan_batch_waves(robot);
    t->crowd_above = crowd_setting(robot);
    return PRRTC_OK;
}

static int solve_batch_call(const char *fn, int n_problems, const prrtc_problem *problems,
                            const prrtc_settings *settings, int n_settings, prrtc_result *results, float *paths,
                            int path_cap, prrtc_batch_stats *stats, bool graph_rounds) {
    return guarded(fn, [&]() -> int {
        if (n_problems < 0 || (n_problems > 0 && (!problems || !settings)) ||
            (n_settings != 1 && n_settings != n_problems)) {
            set_error("%s: invalid argument", fn);
            return PRRTC_EINVAL;
        }
        if (stats) *stats = prrtc_batch_stats{};
        if (n_problems == 0) return PRRTC_OK;
        std::vector<HostProblem> probs;
        for (int i = 0; i < n_problems; i++) {
            results[i] = prrtc_result{};
            probs.push_back(HostProblem{problems[i].robot, problems[i].start, problems[i].goals, problems[i].n_goals,
                                        &problems[i].scene, n_settings == 1 ? settings : settings + i, i});
        }
        return solve_many(probs, results, paths, path_cap, path_cap * kMaxDof, stats, fn, graph_rounds);
    });
}

int prrtc_solve_batch(int n_problems, const prrtc_problem *problems, const prrtc_settings *settings,
                      int n_settings, prrtc_result *results, float *paths, int path_cap,
                      prrtc_batch_stats *stats) {
    return solve_batch_call("prrtc_solve_batch", n_problems, problems, settings, n_settings, results, paths,
                            path_cap, stats, false);
}

int prrtc_solve_batch_rounds(int n_problems, const prrtc_problem *problems, const prrtc_settings *settings,
                             int n_settings, prrtc_result *results, float *paths, int path_cap,
                             prrtc_batch_stats *stats) {
    return solve_batch_call("prrtc_solve_batch_rounds", n_problems, problems, settings, n_settings, results,
                            paths, path_cap, stats, true);
}

}  // extern "C"
