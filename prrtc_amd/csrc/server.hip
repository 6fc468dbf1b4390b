// Continuous-batching planning server over the persistent planner.
//
// The reference answers one blocking solve per process at a time
// (src/planning/pRRTC.cu:607-766).  Here a worker thread keeps a table of
// resident problems per robot in device memory and runs the planner in
// time-sliced persistent launches: between two slices it harvests the
// problems that stopped and admits queued requests into the freed table
// entries and budget.  The planner is round-synchronous and its whole state
// lives in device memory, so where a slice ends changes nothing: a request's
// result equals prrtc_solve's on the same inputs.
//
// Memory: one device allocation of budget_bytes is sub-allocated first-fit
// into per-problem blocks (scene, init rows, trees, Halton states, slot
// outcomes, staged nodes); admission takes requests in arrival order while
// the next one's block fits.  Per robot: a problem table (PlanDesc,
// PlanState) of max_resident entries, a path-row buffer, a stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "planner_host.h"

namespace prrtc {
namespace {

constexpr int kRobots = kNumRobots;

// First-fit sub-allocator over [0, total); sizes are multiples of 256.
class Arena {
public:
    explicit Arena(size_t total) { free_[0] = total; }
    bool alloc(size_t bytes, size_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, o = it->first + bytes;
            free_.erase(it);
            if (rest) free_[o] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;   // offset -> size of free ranges
    size_t used_ = 0;
};

// A request with its own copies of the caller's arrays.
struct Request {
    uint64_t ticket = 0;
    int robot = 0;
    int n_goals = 0;
    prrtc_settings st{};
    std::vector<float> start, goals, tables[5];
    prrtc_scene scene{};
    BlockLayout lay;
    int64_t t_submit = 0, t_admit = 0;
    int64_t kernel_ns = 0;
    int entry = -1;
    size_t off = 0;
    int slices = 0;
    int round = 0;   // planner rounds done, as of the last harvest
    int64_t deadline_ns = 0;   // absolute (now_ns clock), 0: none
    bool cancel = false;       // prrtc_server_cancel asked to end it
    int stop_reason = 0;       // PRRTC_ETIMEDOUT / PRRTC_ECANCELED once its stop word is set
    int stop_slices = 0;       // slices it had run when its stop word was set
    HostProblem hp() const { return HostProblem{robot, start.data(), goals.data(), n_goals, &scene, &st, 0}; }
};

struct Finished {
    int dof = 0;
    prrtc_result r{};
    std::vector<float> path;
    prrtc_request_timing timing{};
};

template <class T>
T *pinned(size_t n) {
    void *p = nullptr;
    PRRTC_HIP(hipHostMalloc(&p, sizeof(T) * std::max<size_t>(n, 1), hipHostMallocDefault));
    return static_cast<T *>(p);
}

// Problem table, buffers and stream of one robot.
struct RobotClass {
    bool ready = false;
    int robot = 0, D = 0, E = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PlanDesc *d_desc = nullptr, *h_desc = nullptr;
    PlanState *d_state = nullptr, *h_state = nullptr;
    PlanControl *d_ctrl = nullptr, *h_ctrl_up = nullptr, *h_ctrl_down = nullptr;
    int *d_list = nullptr, *h_list = nullptr;
    // entry masks, mask_words apart: running | open | aged
    unsigned long long *d_mask = nullptr, *h_mask = nullptr;
    int mask_words = 0;
    float *d_paths = nullptr;
    int path_rows = 0;
    // stop word per table entry (PlanControl::stop_words): set and cleared
    // by the worker thread between slices, in pinned memory, and copied to
    // the device before the next launch when one changed (the planner reads
    // the device copy on every claim and commit)
    int *h_stop = nullptr;
    int *d_stop = nullptr;
    bool stop_dirty = false;
    char *h_stage = nullptr;
    size_t stage_bytes = 0;
    std::vector<Request *> entries;   // resident request per table entry
    std::set<int> free_entries;
    int high = 0;                     // entries [0, high) may be in use
    int running = 0;
    std::vector<Request *> fresh;     // admitted, not yet initialised
    bool launched = false;

    void init(int rb, int max_resident, int path_cap) {
        robot = rb;
        D = robot_dof(rb);
        E = max_resident;
        PRRTC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        PRRTC_HIP(hipEventCreate(&e0));
        PRRTC_HIP(hipEventCreate(&e1));
        PRRTC_HIP(hipMalloc(&d_desc, sizeof(PlanDesc) * E));
        PRRTC_HIP(hipMalloc(&d_state, sizeof(PlanState) * E));
        PRRTC_HIP(hipMalloc(&d_ctrl, sizeof(PlanControl)));
        PRRTC_HIP(hipMalloc(&d_list, sizeof(int) * E));
        mask_words = (E + 63) / 64;
        PRRTC_HIP(hipMalloc(&d_mask, sizeof(unsigned long long) * 3 * mask_words));
        // prrtc_server_create bounds this product by INT32_MAX rows
        path_rows = (int)((int64_t)E * std::max(1, path_cap - 2));
        PRRTC_HIP(hipMalloc(&d_paths, sizeof(float) * (size_t)path_rows * D));
        // status 0 marks a free entry the planner skips
        PRRTC_HIP(hipMemsetAsync(d_state, 0, sizeof(PlanState) * E, stream));
        h_desc = pinned<PlanDesc>(E);
        h_state = pinned<PlanState>(E);
        h_ctrl_up = pinned<PlanControl>(1);
        h_ctrl_down = pinned<PlanControl>(1);
        h_list = pinned<int>(E);
        h_mask = pinned<unsigned long long>(3 * mask_words);
        h_stop = pinned<int>(E);
        std::memset(h_stop, 0, sizeof(int) * (size_t)E);
        PRRTC_HIP(hipMalloc(&d_stop, sizeof(int) * E));
        PRRTC_HIP(hipMemsetAsync(d_stop, 0, sizeof(int) * E, stream));
        entries.assign(E, nullptr);
        for (int e = 0; e < E; e++) free_entries.insert(e);
        PRRTC_HIP(hipStreamSynchronize(stream));
        ready = true;
    }
    void release() {
        if (!ready) return;
        (void)hipStreamSynchronize(stream);
        (void)hipFree(d_desc);
        (void)hipFree(d_state);
        (void)hipFree(d_ctrl);
        (void)hipFree(d_list);
        (void)hipFree(d_mask);
        (void)hipFree(d_paths);
        (void)hipHostFree(h_desc);
        (void)hipHostFree(h_state);
        (void)hipHostFree(h_ctrl_up);
        (void)hipHostFree(h_ctrl_down);
        (void)hipHostFree(h_list);
        (void)hipHostFree(h_mask);
        (void)hipHostFree(h_stop);
        (void)hipFree(d_stop);
        if (h_stage) (void)hipHostFree(h_stage);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(stream);
        ready = false;
    }
    char *stage(size_t n) {
        if (n > stage_bytes) {
            if (h_stage) (void)hipHostFree(h_stage);
            h_stage = nullptr;
            stage_bytes = 0;
            h_stage = pinned<char>(n);
            stage_bytes = n;
        }
        return h_stage;
    }
};

}  // namespace

struct ServerImpl {
    prrtc_server_config cfg{};
    int device = 0;
    unsigned long long ticks_per_us = 100;
    char *arena_dev = nullptr;
    std::unique_ptr<Arena> arena;
    RobotClass cls[kRobots];

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::unique_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> resident;
    std::unordered_map<uint64_t, Finished> done;
    std::unordered_set<uint64_t> live;       // submitted and not yet polled
    uint64_t next_ticket = 1;
    bool stop = false;
    std::string failure;                     // worker's HIP error, if any
    prrtc_server_stats stats{};
    std::thread worker;

    void run();
    void admit_locked();
    void expire_locked(std::vector<std::pair<uint64_t, Finished>> &out);
    void launch(RobotClass &c);
    void harvest(RobotClass &c, std::vector<std::pair<uint64_t, Finished>> &out, std::vector<Request *> &gone,
                 int64_t &kernel_ns);
};

void ServerImpl::admit_locked() {
    int skipped = 0;
    for (auto it = queue.begin(); it != queue.end() && skipped < 256;) {
        Request *rq = it->get();
        RobotClass &c = cls[rq->robot];
        if (!c.ready) c.init(rq->robot, cfg.max_resident, cfg.path_cap);
        if (c.free_entries.empty()) {   // this robot's table is full: later robots may still fit
            ++it;
            ++skipped;
            continue;
        }
        size_t off = 0;
        if (!arena->alloc(rq->lay.bytes, &off)) break;   // arrival order: wait for budget
        rq->off = off;
        rq->entry = *c.free_entries.begin();
        c.free_entries.erase(c.free_entries.begin());
        c.high = std::max(c.high, rq->entry + 1);
        c.entries[rq->entry] = rq;
        if (c.h_stop[rq->entry]) {   // no kernel of any robot runs during admission
            c.h_stop[rq->entry] = 0;
            c.stop_dirty = true;
        }
        c.fresh.push_back(rq);
        rq->t_admit = now_ns();
        resident[rq->ticket] = std::move(*it);
        it = queue.erase(it);
    }
    stats.pending = (int32_t)queue.size();
    stats.resident = (int32_t)resident.size();
    stats.resident_bytes = arena->used();
    stats.peak_resident_bytes = std::max<uint64_t>(stats.peak_resident_bytes, arena->used());
}

// Deadlines and cancellations (worker thread, under the lock, between
// slices): a queued request ends at once, unsolved and never admitted; a
// resident one gets its stop word set, and the planner ends it at its next
// commit (harvest reports it with the stop reason).
void ServerImpl::expire_locked(std::vector<std::pair<uint64_t, Finished>> &out) {
    const int64_t now = now_ns();
    auto reason = [&](const Request &rq) {
        if (rq.cancel) return PRRTC_ECANCELED;
        return rq.deadline_ns && now >= rq.deadline_ns ? PRRTC_ETIMEDOUT : 0;
    };
    for (auto it = queue.begin(); it != queue.end();) {
        Request *rq = it->get();
        const int why = reason(*rq);
        if (!why) { ++it; continue; }
        Finished f;
        f.dof = robot_dof(rq->robot);
        f.r.status = why;
        f.r.wall_ns = now - rq->t_submit;
        f.timing.queued_ns = now - rq->t_submit;
        out.emplace_back(rq->ticket, std::move(f));
        it = queue.erase(it);
    }
    for (auto &kv : resident) {
        Request *rq = kv.second.get();
        if (rq->stop_reason || rq->entry < 0) continue;
        if (const int why = reason(*rq)) {
            rq->stop_reason = why;
            rq->stop_slices = rq->slices;
            cls[rq->robot].h_stop[rq->entry] = 1;
            cls[rq->robot].stop_dirty = true;
        }
    }
    stats.pending = (int32_t)queue.size();
}

// slices a stopped request waits behind young ones before it goes first
constexpr int kStopSlices = 4;

void ServerImpl::launch(RobotClass &c) {
    c.launched = false;
    if (c.running == 0 && c.fresh.empty()) return;
    const hipStream_t s = c.stream;
    if (!c.fresh.empty()) {
        // scene + init rows of each new problem (one copy each), its PlanDesc,
        // then tree roots and Halton states
        size_t need = 0;
        for (Request *rq : c.fresh) need += (rq->lay.upload_bytes + 255) & ~size_t(255);
        char *stg = c.stage(need);
        size_t so = 0;
        int lo = c.E, hi = -1;
        for (size_t i = 0; i < c.fresh.size(); i++) {
            Request *rq = c.fresh[i];
            const HostProblem hp = rq->hp();
            block_upload_image(hp, rq->lay, stg + so);
            PRRTC_HIP(hipMemcpyAsync(arena_dev + rq->off, stg + so, rq->lay.upload_bytes, hipMemcpyHostToDevice, s));
            so += (rq->lay.upload_bytes + 255) & ~size_t(255);
            fill_desc(c.h_desc[rq->entry], hp, rq->lay, arena_dev + rq->off);
            c.h_list[i] = rq->entry;
            lo = std::min(lo, rq->entry);
            hi = std::max(hi, rq->entry);
        }
        PRRTC_HIP(hipMemcpyAsync(c.d_desc + lo, c.h_desc + lo, sizeof(PlanDesc) * (size_t)(hi - lo + 1),
                                 hipMemcpyHostToDevice, s));
        PRRTC_HIP(hipMemcpyAsync(c.d_list, c.h_list, sizeof(int) * c.fresh.size(), hipMemcpyHostToDevice, s));
        launch_plan_init(c.robot, c.d_desc, c.d_state, c.d_list, (int)c.fresh.size(), s);
        c.running += (int)c.fresh.size();
        c.fresh.clear();
    }
    int lds_f4 = 0, ext = 0, max_nnc = 0;
    long long slots = 0;
    const int words = (c.high + 63) / 64;
    const int W = c.mask_words;
    for (int w = 0; w < 3 * W; w++) c.h_mask[w] = 0ull;
    // A stopped request counts as young, which ends it within a few slices
    // while the server keeps up.  Past its capacity the young requests alone
    // can keep a stopped round from being claimed for many slices, so once a
    // stopped request has waited kStopSlices slices every other request is
    // served as aged until it ends: the stopped rounds are claimed first and
    // their results come back near their deadlines.
    bool any_stopped = false;
    for (int e = 0; e < c.high; e++)
        any_stopped |= c.entries[e] && c.entries[e]->stop_reason &&
                       c.entries[e]->slices - c.entries[e]->stop_slices >= kStopSlices;
    for (int e = 0; e < c.high; e++) {
        const Request *rq = c.entries[e];
        if (!rq) continue;
        // every resident entry is running, and open: a launch ends between
        // tasks, so no round is left with all its slots claimed
        c.h_mask[e >> 6] |= 1ull << (e & 63);
        c.h_mask[W + (e >> 6)] |= 1ull << (e & 63);
        // a stopped request counts as young: its slots (skipped tasks) are
        // claimed first, so the round that ends it commits at once instead of
        // waiting behind the young requests of a busy server
        if (!rq->stop_reason && (any_stopped || rq->round >= cfg.young_rounds))
            c.h_mask[2 * W + (e >> 6)] |= 1ull << (e & 63);
        lds_f4 = std::max(lds_f4, (rq->lay.scene_floats - 8) / 4);
        ext = std::max(ext, rq->lay.max_ext);
        max_nnc = std::max(max_nnc, rq->st.num_new_configs);
        slots += rq->st.num_new_configs;
    }
    PlanControl &hc = *c.h_ctrl_up;
    hc = PlanControl{};
    hc.n_problems = c.high;
    hc.path_cap = c.path_rows;
    hc.max_nnc = max_nnc;
    hc.slice_ticks = (unsigned long long)cfg.slice_us * ticks_per_us;
    hc.mask_words = words;
    hc.prio_policy = kPrioYoungest;
    hc.wide_slots = plan_wide_slots();
    hc.crowd_above = plan_crowd_above(c.robot, max_nnc);
    hc.run_mask = c.d_mask;
    hc.open_mask = c.d_mask + W;
    hc.stop_words = c.d_stop;
    if (cfg.young_rounds > 0) {
        hc.aged_mask = c.d_mask + 2 * W;
        hc.young_rounds = cfg.young_rounds;
    }
    PRRTC_HIP(hipMemcpyAsync(c.d_mask, c.h_mask, sizeof(unsigned long long) * 3 * W, hipMemcpyHostToDevice, s));
    if (c.stop_dirty) {
        PRRTC_HIP(hipMemcpyAsync(c.d_stop, c.h_stop, sizeof(int) * (size_t)c.E, hipMemcpyHostToDevice, s));
        c.stop_dirty = false;
    }
    PRRTC_HIP(hipMemcpyAsync(c.d_ctrl, c.h_ctrl_up, sizeof(PlanControl), hipMemcpyHostToDevice, s));
    // throughput shape while many problems are resident, latency shape
    // (obstacles split over the task's waves) for the last few
    const int w_lat = plan_task_waves(c.robot), w_thr = plan_batch_waves(c.robot);
    const int tail = plan_tail_problems(c.robot, max_nnc);
    const int waves = c.running > tail ? w_thr : w_lat;
    const int grid = plan_grid(c.robot, waves, slots);
    PRRTC_HIP(hipEventRecord(c.e0, s));
    launch_plan_slice(c.robot, c.d_desc, c.d_state, c.d_ctrl, c.d_paths, grid, waves, lds_f4, ext, s);
    PRRTC_HIP(hipEventRecord(c.e1, s));
    PRRTC_HIP(hipMemcpyAsync(c.h_state, c.d_state, sizeof(PlanState) * (size_t)c.high, hipMemcpyDeviceToHost, s));
    PRRTC_HIP(hipMemcpyAsync(c.h_ctrl_down, c.d_ctrl, sizeof(PlanControl), hipMemcpyDeviceToHost, s));
    c.launched = true;
}

void ServerImpl::harvest(RobotClass &c, std::vector<std::pair<uint64_t, Finished>> &out,
                         std::vector<Request *> &gone, int64_t &kernel_ns) {
    if (!c.launched) return;
    PRRTC_HIP(hipStreamSynchronize(c.stream));
    float ms = 0.0f;
    PRRTC_HIP(hipEventElapsedTime(&ms, c.e0, c.e1));
    const int64_t slice_ns = (int64_t)(ms * 1e6);
    kernel_ns += slice_ns;
    std::vector<int> stopped;
    bool any_solved = false;
    for (int e = 0; e < c.high; e++) {
        Request *rq = c.entries[e];
        if (!rq) continue;
        rq->slices++;
        rq->kernel_ns += slice_ns;
        const int status = c.h_state[e].status;
        rq->round = c.h_state[e].round;
        if (status == kStatusRunning) continue;
        stopped.push_back(e);
        any_solved |= status == kStatusSolved;
    }
    if (stopped.empty()) return;
    std::vector<float> rows;
    if (any_solved) {
        const int used = std::min(c.h_ctrl_down->path_rows, c.h_ctrl_down->path_cap);
        rows.resize((size_t)used * c.D);
        if (used > 0)
            PRRTC_HIP(hipMemcpy(rows.data(), c.d_paths, sizeof(float) * rows.size(), hipMemcpyDeviceToHost));
    }
    const int64_t t_done = now_ns();
    for (int e : stopped) {
        Request *rq = c.entries[e];
        const PlanState &ps = c.h_state[e];
        Finished f;
        f.dof = c.D;
        prrtc_result &r = f.r;
        r.start_tree_size = ps.size[0];
        r.goal_tree_size = ps.size[1];
        r.edge_checks = (int64_t)ps.edge_checks;
        r.pose_checks = (int64_t)ps.pose_checks;
        r.solved = ps.status == kStatusSolved;
        r.status = r.solved                          ? PRRTC_OK
                   : ps.status == kStatusPathLimit ? PRRTC_EOVERFLOW
                   : ps.status == kStatusStopped   ? rq->stop_reason
                                                   : PRRTC_OK;
        r.iters = r.solved ? ps.iters : 0;
        r.cost = r.solved ? ps.cost : 0.0f;
        if (r.solved) {
            f.path.resize((size_t)cfg.path_cap * c.D);
            const int n = assemble_path(rq->hp(), ps, rows.data(), c.D, cfg.path_cap, f.path.data());
            if (n < 0) {
                r.solved = 0;
                r.status = PRRTC_EOVERFLOW;
                f.path.clear();
            } else {
                r.path_nodes = n;
                r.path_length = ps.t_len + ps.o_len;
                f.path.resize((size_t)n * c.D);
            }
        }
        r.wall_ns = t_done - rq->t_submit;
        r.kernel_ns = rq->kernel_ns;
        f.timing.queued_ns = rq->t_admit - rq->t_submit;
        f.timing.resident_ns = t_done - rq->t_admit;
        f.timing.slices = rq->slices;
        c.entries[e] = nullptr;
        c.free_entries.insert(e);
        c.running--;
        out.emplace_back(rq->ticket, std::move(f));
        gone.push_back(rq);
    }
    while (c.high > 0 && !c.entries[c.high - 1]) c.high--;
}

void ServerImpl::run() {
    try {
        PRRTC_HIP(hipSetDevice(device));
        while (true) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !queue.empty() || !resident.empty(); });
                if (stop) break;
                std::vector<std::pair<uint64_t, Finished>> dropped;
                expire_locked(dropped);
                for (auto &kv : dropped) {
                    stats.completed++;
                    stats.timed_out += kv.second.r.status == PRRTC_ETIMEDOUT;
                    stats.cancelled += kv.second.r.status == PRRTC_ECANCELED;
                    done[kv.first] = std::move(kv.second);
                }
                if (!dropped.empty()) cv_done.notify_all();
                admit_locked();
            }
            for (RobotClass &c : cls)
                if (c.ready) launch(c);
            std::vector<std::pair<uint64_t, Finished>> out;
            std::vector<Request *> gone;
            int64_t slices = 0, kernel_ns = 0;
            for (RobotClass &c : cls) {
                if (!c.ready || !c.launched) continue;
                slices++;
                harvest(c, out, gone, kernel_ns);
            }
            std::lock_guard<std::mutex> lk(mu);
            stats.slices += slices;
            stats.kernel_ns += kernel_ns;
            for (Request *rq : gone) arena->release(rq->off, rq->lay.bytes);
            for (auto &kv : out) {
                stats.completed++;
                stats.solved += kv.second.r.solved;
                stats.timed_out += kv.second.r.status == PRRTC_ETIMEDOUT;
                stats.cancelled += kv.second.r.status == PRRTC_ECANCELED;
                stats.edge_checks += kv.second.r.edge_checks;
                done[kv.first] = std::move(kv.second);
                resident.erase(kv.first);
            }
            stats.resident = (int32_t)resident.size();
            stats.resident_bytes = arena->used();
            if (!out.empty()) cv_done.notify_all();
        }
    } catch (const HipError &e) {
        std::lock_guard<std::mutex> lk(mu);
        char buf[256];
        std::snprintf(buf, sizeof(buf), "%s failed: %s", e.where, hipGetErrorString(e.code));
        failure = buf;
        cv_done.notify_all();
    } catch (const std::bad_alloc &) {
        std::lock_guard<std::mutex> lk(mu);
        failure = "host allocation failed";
        cv_done.notify_all();
    }
}

}  // namespace prrtc

struct prrtc_server {
    prrtc::ServerImpl impl;
};

using namespace prrtc;

extern "C" {

void prrtc_server_config_default(prrtc_server_config *cfg) {
    if (!cfg) return;
    *cfg = prrtc_server_config{};
    cfg->budget_bytes = 2ull << 30;
    cfg->max_resident = 1024;
    cfg->max_pending = 65536;
    cfg->slice_us = 1000;
    cfg->path_cap = 1024;
    cfg->young_rounds = kYoungRounds;
    cfg->deadline_ms = 0;
}

int prrtc_server_create(const prrtc_server_config *cfg, prrtc_server **out) {
    return guarded("prrtc_server_create", [&]() -> int {
        if (!out) { set_error("prrtc_server_create: out is NULL"); return PRRTC_EINVAL; }
        *out = nullptr;
        prrtc_server_config c;
        prrtc_server_config_default(&c);
        if (cfg) {
            if (cfg->budget_bytes) c.budget_bytes = cfg->budget_bytes;
            if (cfg->max_resident) c.max_resident = cfg->max_resident;
            if (cfg->max_pending) c.max_pending = cfg->max_pending;
            if (cfg->slice_us) c.slice_us = cfg->slice_us;
            if (cfg->path_cap) c.path_cap = cfg->path_cap;
            if (cfg->young_rounds) c.young_rounds = cfg->young_rounds;
            c.deadline_ms = cfg->deadline_ms;
        }
        if (c.max_resident < 1 || c.max_resident > (1 << 20) || c.max_pending < 1 || c.slice_us < 1 ||
            c.path_cap < 2 || c.budget_bytes < 4096 || c.deadline_ms < 0) {
            set_error("prrtc_server_create: invalid config");
            return PRRTC_EINVAL;
        }
        // each robot class keeps path_cap - 2 rows per resident entry; the
        // pool is indexed with 32-bit row numbers on the device
        const int64_t pool_rows = (int64_t)c.max_resident * std::max(1, c.path_cap - 2);
        if (pool_rows > INT32_MAX) {
            set_error("prrtc_server_create: max_resident x (path_cap - 2) = %lld path rows exceeds %d",
                      (long long)pool_rows, INT32_MAX);
            return PRRTC_EINVAL;
        }
        auto srv = std::make_unique<prrtc_server>();
        ServerImpl &S = srv->impl;
        S.cfg = c;
        PRRTC_HIP(hipGetDevice(&S.device));
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, S.device) == hipSuccess && khz > 0)
            S.ticks_per_us = (unsigned long long)std::max(1, khz / 1000);
        const size_t budget = (size_t)c.budget_bytes & ~size_t(255);
        if (hipMalloc(&S.arena_dev, budget) != hipSuccess) {
            (void)hipGetLastError();
            set_error("prrtc_server_create: cannot reserve %zu bytes of device memory", budget);
            return PRRTC_ENOMEM;
        }
        S.arena = std::make_unique<Arena>(budget);
        S.worker = std::thread([&S] { S.run(); });
        *out = srv.release();
        return PRRTC_OK;
    });
}

int prrtc_server_submit(prrtc_server *srv, const prrtc_problem *problem, const prrtc_settings *settings,
                        uint64_t *ticket) {
    return guarded("prrtc_server_submit", [&]() -> int {
        if (!srv || !problem || !ticket) { set_error("prrtc_server_submit: NULL argument"); return PRRTC_EINVAL; }
        *ticket = 0;
        ServerImpl &S = srv->impl;
        const HostProblem in{problem->robot, problem->start, problem->goals, problem->n_goals, &problem->scene,
                             settings, 0};
        if (int rc = check_problem(in, "prrtc_server_submit")) return rc;
        auto rq = std::make_unique<Request>();
        const int D = robot_dof(in.robot);
        rq->robot = in.robot;
        rq->n_goals = in.n_goals;
        rq->st = *settings;
        rq->start.assign(in.start, in.start + D);
        rq->goals.assign(in.goals, in.goals + (size_t)D * in.n_goals);
        const prrtc_scene &sc = problem->scene;
        const float *src[5] = {sc.spheres, sc.capsules, sc.z_capsules, sc.cuboids, sc.z_cuboids};
        const int cnt[5] = {sc.n_spheres, sc.n_capsules, sc.n_z_capsules, sc.n_cuboids, sc.n_z_cuboids};
        const int width[5] = {4, 8, 8, 15, 15};
        for (int k = 0; k < 5; k++)
            if (cnt[k] > 0) rq->tables[k].assign(src[k], src[k] + (size_t)cnt[k] * width[k]);
        prrtc_scene &own = rq->scene;
        own.spheres = rq->tables[0].data();    own.n_spheres = cnt[0];
        own.capsules = rq->tables[1].data();   own.n_capsules = cnt[1];
        own.z_capsules = rq->tables[2].data(); own.n_z_capsules = cnt[2];
        own.cuboids = rq->tables[3].data();    own.n_cuboids = cnt[3];
        own.z_cuboids = rq->tables[4].data();  own.n_z_cuboids = cnt[4];
        rq->lay = block_layout(rq->hp());
        std::lock_guard<std::mutex> lk(S.mu);
        if (!S.failure.empty()) { set_error("prrtc_server_submit: worker failed: %s", S.failure.c_str()); return PRRTC_EDEVICE; }
        if (rq->lay.bytes > (size_t)S.cfg.budget_bytes) {
            set_error("prrtc_server_submit: problem needs %zu bytes, budget is %llu", rq->lay.bytes,
                      (unsigned long long)S.cfg.budget_bytes);
            return PRRTC_ENOMEM;
        }
        if ((int)S.queue.size() >= S.cfg.max_pending) { set_error("prrtc_server_submit: queue full"); return PRRTC_EBUSY; }
        rq->ticket = S.next_ticket++;
        rq->t_submit = now_ns();
        if (S.cfg.deadline_ms > 0) rq->deadline_ns = rq->t_submit + (int64_t)S.cfg.deadline_ms * 1000000;
        *ticket = rq->ticket;
        S.live.insert(rq->ticket);
        S.queue.push_back(std::move(rq));
        S.stats.submitted++;
        S.stats.pending = (int32_t)S.queue.size();
        S.cv_work.notify_one();
        return PRRTC_OK;
    });
}

int prrtc_server_poll(prrtc_server *srv, uint64_t ticket, int32_t wait_ms, prrtc_result *result, float *path,
                      int32_t path_cap, prrtc_request_timing *timing) {
    return guarded("prrtc_server_poll", [&]() -> int {
        if (!srv || !result) { set_error("prrtc_server_poll: NULL argument"); return PRRTC_EINVAL; }
        ServerImpl &S = srv->impl;
        std::unique_lock<std::mutex> lk(S.mu);
        if (!S.live.count(ticket)) { set_error("prrtc_server_poll: unknown ticket %llu", (unsigned long long)ticket); return PRRTC_EINVAL; }
        auto ready = [&] { return S.done.count(ticket) || !S.failure.empty(); };
        if (wait_ms < 0) S.cv_done.wait(lk, ready);
        else S.cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
        auto it = S.done.find(ticket);
        if (it == S.done.end()) {
            if (!S.failure.empty()) { set_error("prrtc_server_poll: worker failed: %s", S.failure.c_str()); return PRRTC_EDEVICE; }
            set_error("prrtc_server_poll: ticket %llu not finished", (unsigned long long)ticket);
            return PRRTC_EAGAIN;
        }
        const Finished &f = it->second;
        const int D = f.dof;
        if (f.r.path_nodes > 0 && (!path || path_cap < f.r.path_nodes)) {
            set_error("prrtc_server_poll: path needs %d rows, path_cap is %d", f.r.path_nodes, path_cap);
            return PRRTC_EOVERFLOW;   // the result stays available
        }
        *result = f.r;
        if (f.r.path_nodes > 0) std::memcpy(path, f.path.data(), sizeof(float) * (size_t)f.r.path_nodes * D);
        if (timing) *timing = f.timing;
        S.done.erase(it);
        S.live.erase(ticket);
        return PRRTC_OK;
    });
}

int prrtc_server_get_stats(prrtc_server *srv, prrtc_server_stats *out) {
    if (!srv || !out) { set_error("prrtc_server_get_stats: NULL argument"); return PRRTC_EINVAL; }
    std::lock_guard<std::mutex> lk(srv->impl.mu);
    *out = srv->impl.stats;
    return PRRTC_OK;
}

int prrtc_server_cancel(prrtc_server *srv, uint64_t ticket) {
    return guarded("prrtc_server_cancel", [&]() -> int {
        if (!srv) { set_error("prrtc_server_cancel: NULL server"); return PRRTC_EINVAL; }
        ServerImpl &S = srv->impl;
        std::lock_guard<std::mutex> lk(S.mu);
        if (!S.live.count(ticket)) {
            set_error("prrtc_server_cancel: unknown ticket %llu", (unsigned long long)ticket);
            return PRRTC_EINVAL;
        }
        if (S.done.count(ticket)) return PRRTC_OK;   // finished: the result stands
        // the worker applies it at its next pass (it owns the stop words)
        for (auto &q : S.queue)
            if (q->ticket == ticket) q->cancel = true;
        auto it = S.resident.find(ticket);
        if (it != S.resident.end()) it->second->cancel = true;
        S.cv_work.notify_one();
        return PRRTC_OK;
    });
}

int prrtc_server_destroy(prrtc_server *srv) {
    if (!srv) return PRRTC_OK;
    ServerImpl &S = srv->impl;
    {
        std::lock_guard<std::mutex> lk(S.mu);
        S.stop = true;
    }
    S.cv_work.notify_all();
    if (S.worker.joinable()) S.worker.join();
    for (RobotClass &c : S.cls) c.release();
    if (S.arena_dev) (void)hipFree(S.arena_dev);
    delete srv;
    return PRRTC_OK;
}

}  // extern "C"
