// Round-synchronous pRRTC planner on gfx950: device side.
//
// The reference runs num_new_configs (nnc) CUDA blocks per problem, each
// looping sample -> nearest -> steer -> validate -> insert -> connect with
// the trees shared through global atomics (src/planning/pRRTC.cu:305-602).
// Here the same per-sample step is a *task* (problem p, slot s) executed by
// one workgroup of 1-4 wave64s, and a problem advances in rounds: every slot
// of a round works against the trees as committed at the round start; the
// workgroup that finishes the round's last task commits the new nodes in
// slot order, applies the dynamic-domain radius updates in slot order, and
// detects the solution.  Results are therefore a deterministic function of
// (problem, settings, seed), identical whether a problem runs alone or among
// hundreds, and whatever the launch shape.
//
// Persistent kernel: a fixed grid of task workgroups pulls tasks of any
// running problem from per-problem slot counters; problems of one launch
// have independent trees, scenes and settings.
#pragma once

#include "device_common.h"

namespace prrtc {

constexpr int kStatusRunning = 1;
constexpr int kStatusSolved = 2;
constexpr int kStatusIterLimit = -1;     // max_iters rounds without a solution
constexpr int kStatusNodeLimit = -2;     // a tree would exceed max_samples
constexpr int kStatusPathLimit = -3;     // solution path longer than the buffer
constexpr int kStatusStopped = -4;       // stopped by the host (deadline or cancel)
constexpr int kClaimedOut = 1 << 30;     // next_slot of a finished problem
constexpr int kMaxExt = 255;             // longest connect chain (sub-edges)
constexpr int kMaxNnc = 1024;

// Per-problem description: read-only during the launch.
struct PlanDesc {
    int nnc, g, max_iters, max_samples;
    int balance, dd, n_goals, cap;
    int max_ext, scene_f4;
    int robot;            // PRRTC_ROBOT_*: the planner code a mixed launch dispatches to
    int wide_from;        // chain pass from which a task's waves take distinct poses (< 0: never)
    float range, tree_ratio, dd_alpha, dd_radius;
    float dd_min_radius;
    int wide_edge;        // a steered edge of more than 64 poses spreads them over pose groups (0: never)
    float pad3, pad4;
    const float *scene;   // packed scene
    float *nodes;         // [2][D][cap]   (SoA per tree)
    int *parent;          // [2][cap]
    float *radii;         // [2][cap]
    float *halton;        // [nnc][3][D]   bases, numerators, denominators
    int *tprev;           // [nnc]         tree each slot extended last round
    int *meta;            // [nnc][8]      per-slot outcome of the round
    float *stage;         // [nnc][1+max_ext][D]  new node + extension nodes
    const float *init;    // start [D] then goals [n_goals][D]
    uint64_t seed;
};

// Per-problem mutable state.
struct PlanState {
    int status;
    int round;            // 1-based round (= the reference's per-block iter)
    int next_slot;        // slot claim counter of the current round
    int done_slots;       // finished tasks of the current round
    int size[2];          // committed tree sizes
    int iters;            // round that found the solution
    int t_tree;           // tree the connecting slot extended
    int t_len, o_len;     // path rows traced in each tree
    int reached;          // goal index reached
    int path_off;         // first row of this problem's path in the output
    float cost;
    int pad0;
    unsigned long long edge_checks;
    unsigned long long pose_checks;   // lanes carrying a pose at the start of a pass, summed over passes
};

// Launch-wide control block.
struct PlanControl {
    int n_problems;            // table size (running or not)
    int path_rows;             // rows used in the path output
    int path_cap;              // rows available in the path output
    int max_nnc;               // largest num_new_configs of the launch
    int mask_words;            // ceil(n_problems / 64)
    int prio_policy;           // wave priority of a task: kPrioNone / kPrioOldest / kPrioYoungest
    // floats between path rows in the path output: the largest DOF of a
    // launch whose problems have different robots (0: the robot's DOF)
    int path_stride;
    // a task's connect-chain passes may give each wave poses of its own
    // while (running problems) x max_nnc is at most this (0: never): the
    // rounds are then bound by their slowest task, not by throughput
    int wide_slots;
    unsigned long long t0;          // launch start (s_memrealtime ticks), set by the first workgroup
    unsigned long long slice_ticks; // 0: run until every problem stops; else leave after this long
    unsigned long long *run_mask;   // bit p set while problem p runs; cleared by its last commit
    // bit p set while problem p's current round has unclaimed slots: cleared
    // by the claim of its last slot, set again by the round's commit
    unsigned long long *open_mask;
    // young-first claims (a server): bit p set once problem p has done
    // young_rounds rounds, by the commit that reaches it.  Open problems
    // without the bit are served first; null: no preference
    unsigned long long *aged_mask;
    int young_rounds;
    // more running problems than this: a workgroup claims one slot per wave
    // and each wave runs its own (0: always one slot on all waves)
    int crowd_above;
    // the running problem furthest along, (round << 32) | p, or 0: its open
    // slots are claimed first and its tasks run at raised wave priority, so
    // the longest problem of a batch is not slowed by the crowd
    unsigned long long hot_key;
    // host-set stop word per problem (null: none): once nonzero, the
    // problem's tasks are skipped as they are claimed and its next commit
    // ends it with kStatusStopped (a solution found in that round still wins)
    const int *stop_words;
};

__device__ __forceinline__ bool stop_requested(const PlanControl *ctrl, int p) {
    return ctrl->stop_words &&
           __hip_atomic_load(ctrl->stop_words + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// slot outcome flags in meta[1]
constexpr int kSkip = 1, kExtOk = 2, kConnected = 4;

// Task wave priority (s_setprio) from the problem's age in rounds: a batch
// favours its oldest problems (they bound the batch's makespan), a server
// its youngest (they bound request latency).
constexpr int kPrioNone = 0, kPrioOldest = 1, kPrioYoungest = 2;
// a server claims the open slots of problems below this many rounds first
// (its default; the server config sets it)
constexpr int kYoungRounds = 64;
__device__ __forceinline__ int task_priority(int policy, int round, bool hot) {
    const int age = round < 8 ? 0 : round < 64 ? 1 : round < 512 ? 2 : 3;
    if (policy == kPrioOldest) return hot ? 3 : (age < 2 ? age : 2);
    if (policy == kPrioYoungest) return 3 - age;
    return 0;
}

// A planner task runs on one workgroup of 1..kMaxTaskWaves waves.
constexpr int kMaxTaskWaves = 4;
// Nearest-node scans read trees in chunks of this many nodes; tree storage
// is a whole number of chunks.
constexpr int kScanChunk = 1024;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }

__device__ __forceinline__ int atomic_add_agent(int *p, int v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int load_acquire(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_release(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane-0 atomic, result broadcast to the wave.
__device__ __forceinline__ int wave_atomic_add(int *p, int v) {
    int r = 0;
    if (lane_id() == 0) r = atomic_add_agent(p, v);
    return __builtin_amdgcn_readfirstlane(r);
}

// ----------------------------------------------------------- planner math
// All arithmetic below is pinned (contraction off, explicit fmaf where the
// reference's CUDA build fuses); the CPU oracle's planner uses the same.

// Tree choice of a slot (pRRTC.cu:348-371).
__device__ __forceinline__ int choose_tree(const PlanDesc &d, int slot, int iter, int t_prev, int n0, int n1) {
#pragma clang fp contract(off)
    int t = t_prev;
    const int nnc = d.nnc;
    if (d.balance == 0 || iter == 1) {
        t = (slot < (nnc / 2)) ? 0 : 1;
    } else if (d.balance == 1 && abs(n0 - n1) < 1.5 * nnc) {
        const float ratio = n0 / (float)(n0 + n1);
        const float bf = 1 - ratio;
        t = (slot < (nnc * bf)) ? 0 : 1;
    } else if (d.balance == 1) {
        const float ratio = n0 / (float)(n0 + n1);
        t = (ratio < d.tree_ratio) ? 0 : 1;
    } else if (d.balance == 2) {
        const int nt = t == 0 ? n0 : n1, no = t == 0 ? n1 : n0;
        const float ratio = abs(nt - no) / (float)nt;
        if (ratio < d.tree_ratio) t = 1 - t;
    }
    return t;
}

// Incremental Halton step of one dimension (pRRTC.cu:81-105); uniform.
__device__ __forceinline__ float halton_step(float &b, float &n, float &dd) {
#pragma clang fp contract(off)
    const float xf = dd - n;
    if (xf == 1.0f) {
        dd = floorf(dd * b);
        n = 1.0f;
    } else {
        float y = floorf(dd / b);
        // y shrinks by a factor b >= 3 per step; 64 steps bound any float state
        for (int guard = 0; xf <= y && guard < 64; guard++) y = floorf(y / b);
        n = floorf((b + 1.0f) * y) - xf;
    }
    return n / dd;
}

// SplitMix64 (Steele, Lea & Flood 2014).
__host__ __device__ inline uint64_t splitmix64(uint64_t &st) {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Bases of a slot: the first D of {3, 5, ..., 59}, Fisher-Yates shuffled
// unless slot 0 (pRRTC.cu:49-65); the shuffle draws from SplitMix64 seeded
// with seed * phi64 + slot * 0xD1B54A32D192ED03.
__host__ __device__ inline void halton_bases(uint64_t seed, int slot, int dof, float *out) {
    float primes[16] = {3.f, 5.f, 7.f, 11.f, 13.f, 17.f, 19.f, 23.f, 29.f, 31.f, 37.f, 41.f, 43.f, 47.f, 53.f, 59.f};
    if (slot != 0) {
        uint64_t st = seed * 0x9E3779B97F4A7C15ULL + (uint64_t)slot * 0xD1B54A32D192ED03ULL;
        for (int i = 15; i > 0; i--) {
            const int j = (int)(splitmix64(st) % (uint64_t)(i + 1));
            const float tmp = primes[i];
            primes[i] = primes[j];
            primes[j] = tmp;
        }
    }
    for (int i = 0; i < dof; i++) out[i] = primes[i];
}

// One pass over a launch's problem masks, one word per lane (64 words per
// step): the number of running problems, and the first problem at or after
// `from` (cyclically) that runs, the first whose current round has
// unclaimed slots, and the first young one among those (none without an
// aged mask).  Wave-uniform results.
struct MaskScan {
    int running;      // set bits of the running mask
    int first_run;    // first running problem from `from` on, or -1
    int open;         // first open problem from `from` on, or -1
    int open_young;   // first open problem without the aged bit, or -1
    bool want_open;   // the problem asked about has an open round
};

struct FirstFrom {   // first set bit at or after a cursor, wrapping around
    int a = -1, b = -1;
    __device__ __forceinline__ void step(unsigned long long m, int w, int w0, unsigned long long hi_part, int base) {
        const unsigned long long ma = w > w0 ? m : (w == w0 ? m & hi_part : 0ull);
        const unsigned long long mb = w < w0 ? m : (w == w0 ? m & ~hi_part : 0ull);
        const uint64_t ba = ballot(ma != 0ull), bb = ballot(mb != 0ull);
        if (b < 0 && bb) b = first_bit(mb, bb, base);
        if (a < 0 && ba) a = first_bit(ma, ba, base);
    }
    __device__ __forceinline__ int get() const { return a >= 0 ? a : b; }
    // bit index of the first set bit of the word held by lane ctz(lanes)
    __device__ __forceinline__ static int first_bit(unsigned long long m, uint64_t lanes, int base) {
        const int j = __builtin_ctzll(lanes);
        const unsigned lo = __builtin_amdgcn_readlane((unsigned)m, j);
        const unsigned hi = __builtin_amdgcn_readlane((unsigned)(m >> 32), j);
        return (base + j) * 64 + (lo ? __builtin_ctz(lo) : 32 + __builtin_ctz(hi));
    }
};

// `want` (-1: none) is a problem whose open bit is reported in want_open.
__device__ __forceinline__ MaskScan scan_masks(const PlanControl *ctrl, int words, int from, int want) {
    const int lane = lane_id();
    const int w0 = from >> 6;
    const unsigned long long hi_part = ~0ull << (from & 63);
    const unsigned long long *aged = ctrl->aged_mask;
    FirstFrom fr, fo, fy;
    int bits = 0;
    bool want_open = false;
    for (int base = 0; base < words; base += kWave) {
        const int w = base + lane;
        unsigned long long r = 0ull, o = 0ull, a = 0ull;
        if (w < words) {
            r = __hip_atomic_load(ctrl->run_mask + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            o = __hip_atomic_load(ctrl->open_mask + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (aged) a = __hip_atomic_load(aged + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        bits += __builtin_popcountll(r);
        fr.step(r, w, w0, hi_part, base);
        o &= r;
        want_open |= ballot(want >= 0 && w == (want >> 6) && ((o >> (want & 63)) & 1ull)) != 0;
        fo.step(o, w, w0, hi_part, base);
        if (aged) fy.step(o & ~a, w, w0, hi_part, base);
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) bits += __shfl_xor(bits, off);
    MaskScan ms;
    ms.running = __builtin_amdgcn_readfirstlane(bits);
    ms.first_run = fr.get();
    ms.open = fo.get();
    ms.open_young = fy.get();
    ms.want_open = want_open;
    return ms;
}

// Nearest node of tree t to q among [0, n): squared L2 as an fmaf chain in
// dimension order (utils.cuh:100-109), ties to the smaller index.  The
// workgroup's waves stride the SoA coordinates (coalesced) and reduce by lane
// shuffles, then across waves through `red` (LDS, 2 * kMaxTaskWaves words).
// `tree` and `cap` keep every column 16-byte aligned (cap % 4 == 0).  Every
// thread of the workgroup calls it with the same arguments.
// Floats of one task's LDS area: chain rows (1 + max_ext configurations),
// 4 uniform vectors, and the cross-wave reduction buffers (2 x 2 x
// kMaxTaskWaves words).  The leader's decision takes kCtlInts words.
__host__ __device__ constexpr int task_area_floats(int D, int max_ext) {
    return (max_ext + 5) * D + 4 * kMaxTaskWaves;
}
constexpr int kCtlInts = 8;

// The threads that run one planner task: the whole workgroup (its waves
// share the work and meet at workgroup barriers) or a single wave on its
// own (then no workgroup barrier may be used: the other waves run other
// tasks).  `waves` is uniform over the workgroup.
struct Team {
    int waves;   // 1..kMaxTaskWaves
    int wave;    // this wave's index in the team
    int nt;      // threads: kWave * waves
    int tid;     // this thread's index in the team
    __device__ __forceinline__ void sync() const {
        if (waves > 1) {
            __syncthreads();
        } else {   // one wave: order its LDS accesses, no workgroup barrier
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
};
__device__ __forceinline__ Team team_workgroup() {
    return Team{(int)(blockDim.x / kWave), wave_id(), (int)blockDim.x, (int)threadIdx.x};
}
__device__ __forceinline__ Team team_wave() { return Team{1, 0, kWave, lane_id()}; }

template <int D>
__device__ __forceinline__ int nearest(const float *tree, int cap, int n, const float (&q)[D], float &best_sq,
                                       float *red, const Team &tm) {
#pragma clang fp contract(off)
    // (best, idx) starts at (+inf, 0): a finite distance always wins, and a
    // query whose distances all overflow to +inf (or are NaN) still gets a
    // node of the tree, never an index past it
    float best = __builtin_huge_valf();
    int idx = 0;
    const int lane = lane_id();
    const int W = tm.waves, wv = tm.wave;
    // kScanChunk-node chunks, chunk k to wave k % W: lane l reads nodes
    // c + 256u + 4l .. +3 (u = 0..3) of every column as float4, four 16-byte
    // loads per coordinate (the compiled loop waits for each coordinate's
    // loads before the next coordinate's).  The last chunk is read whole too
    // (the tree's capacity is a multiple of kScanChunk) and nodes >= n are
    // masked.
    // Each lane visits its nodes in increasing index order.
    static_assert(kScanChunk == 16 * kWave, "scan chunk = 16 nodes per lane");
    for (int c = wv * kScanChunk; c < n; c += W * kScanChunk) {
        float acc[16];
#pragma unroll
        for (int e = 0; e < 16; e++) acc[e] = 0.0f;
#pragma unroll
        for (int d = 0; d < D; d++) {
            const float4 *col = reinterpret_cast<const float4 *>(tree + (size_t)d * cap + c) + lane;
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = col[u * kWave];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                float diff = v[u].x - q[d];
                acc[4 * u + 0] = fmaf(diff, diff, acc[4 * u + 0]);
                diff = v[u].y - q[d];
                acc[4 * u + 1] = fmaf(diff, diff, acc[4 * u + 1]);
                diff = v[u].z - q[d];
                acc[4 * u + 2] = fmaf(diff, diff, acc[4 * u + 2]);
                diff = v[u].w - q[d];
                acc[4 * u + 3] = fmaf(diff, diff, acc[4 * u + 3]);
            }
        }
#pragma unroll
        for (int e = 0; e < 16; e++) {
            const int i = c + (e >> 2) * 4 * kWave + 4 * lane + (e & 3);
            if (i < n && acc[e] < best) { best = acc[e]; idx = i; }
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const float ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(idx, off);
        if (ob < best || (ob == best && oi < idx)) { best = ob; idx = oi; }
    }
    if (W > 1) {
        if (lane == 0) {
            red[2 * wv] = best;
            red[2 * wv + 1] = __int_as_float(idx);
        }
        tm.sync();
        best = red[0];
        idx = __float_as_int(red[1]);
        for (int w = 1; w < W; w++) {
            const float ob = red[2 * w];
            const int oi = __float_as_int(red[2 * w + 1]);
            if (ob < best || (ob == best && oi < idx)) { best = ob; idx = oi; }
        }
        tm.sync();   // `red` may be written again right after
    }
    best_sq = best;
    return idx;
}

}  // namespace prrtc
