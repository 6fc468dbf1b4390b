// Drop-in for pRRTC's src/planning/Planners.hh: PlannerResult and
// pRRTC::solve<Robot> (Planners.hh:11-23, 75-78), implemented over the C ABI
// of libprrtc.so (include/prrtc.h) -- the MI355X planner.
//
//   auto res = pRRTC::solve<ppln::robots::Panda>(start, goals, env, settings);
//
// Same arguments and result as the reference: the path runs goal -> start
// with the junction node from both trees, path_length counts the interior
// nodes, cost is the sum of edge lengths.  Unlike the reference the call is
// re-entrant, never resets the device, and reports failures (bad arguments,
// device errors) on stderr with solved = false instead of exiting.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <vector>

#include "../../../prrtc.h"
#include "../collision/environment.hh"
#include "Robots.hh"
#include "pRRTC_settings.hh"

template <typename Robot>
struct PlannerResult {
    bool solved = false;
    std::vector<typename Robot::Configuration> path;
    int start_tree_size = 0;
    int goal_tree_size = 0;
    int path_length = 0;
    int iters = 0;
    float cost = 0.0;
    std::size_t wall_ns = 0;    // the whole solve call
    std::size_t kernel_ns = 0;  // device time of the planner
    std::size_t copy_ns = 0;    // host <-> device copies
};

template <typename Robot>
inline float l2dist(typename Robot::Configuration &a, typename Robot::Configuration &b) {
    float s = 0;
    for (std::size_t i = 0; i < a.size(); i++) s += (a[i] - b[i]) * (a[i] - b[i]);
    return std::sqrt(s);
}

template <typename Robot>
inline void print_cfg_ptr(float *config) {
    for (int i = 0; i < Robot::dimension; i++) std::cout << config[i] << " ";
    std::cout << "\n";
}

template <typename Robot>
inline void print_cfg(typename Robot::Configuration &config) {
    print_cfg_ptr<Robot>(config.data());
}

inline std::size_t get_elapsed_nanoseconds(const std::chrono::time_point<std::chrono::steady_clock> &start) {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - start).count();
}

namespace prrtc_detail {

// Row-major obstacle tables of an Environment, in the C ABI layout.
struct SceneTables {
    std::vector<float> spheres, capsules, z_capsules, cuboids, z_cuboids;
    prrtc_scene scene{};
};

template <typename Caps>
inline void put_capsules(std::vector<float> &out, const Caps *c, unsigned n) {
    for (unsigned i = 0; i < n; i++)
        out.insert(out.end(), {c[i].x1, c[i].y1, c[i].z1, c[i].xv, c[i].yv, c[i].zv, c[i].r, c[i].rdv});
}

template <typename Cub>
inline void put_cuboids(std::vector<float> &out, const Cub *c, unsigned n) {
    for (unsigned i = 0; i < n; i++)
        out.insert(out.end(), {c[i].x, c[i].y, c[i].z, c[i].axis_1_x, c[i].axis_1_y, c[i].axis_1_z, c[i].axis_2_x,
                               c[i].axis_2_y, c[i].axis_2_z, c[i].axis_3_x, c[i].axis_3_y, c[i].axis_3_z, c[i].axis_1_r,
                               c[i].axis_2_r, c[i].axis_3_r});
}

inline void tables(const ppln::collision::Environment<float> &env, SceneTables &t) {
    for (unsigned i = 0; i < env.num_spheres; i++)
        t.spheres.insert(t.spheres.end(), {env.spheres[i].x, env.spheres[i].y, env.spheres[i].z, env.spheres[i].r});
    put_capsules(t.capsules, env.capsules, env.num_capsules);
    put_capsules(t.z_capsules, env.z_aligned_capsules, env.num_z_aligned_capsules);
    put_cuboids(t.cuboids, env.cuboids, env.num_cuboids);
    put_cuboids(t.z_cuboids, env.z_aligned_cuboids, env.num_z_aligned_cuboids);
    // env.cylinders is not tested, as in the reference checker
    t.scene = prrtc_scene{t.spheres.data(), (int32_t)env.num_spheres, t.capsules.data(), (int32_t)env.num_capsules,
                          t.z_capsules.data(), (int32_t)env.num_z_aligned_capsules, t.cuboids.data(),
                          (int32_t)env.num_cuboids, t.z_cuboids.data(), (int32_t)env.num_z_aligned_cuboids};
}

inline prrtc_settings c_settings(const pRRTC_settings &s) {
    prrtc_settings c;
    c.max_samples = s.max_samples;
    c.max_iters = s.max_iters;
    c.num_new_configs = s.num_new_configs;
    c.granularity = s.granularity;
    c.range = s.range;
    c.balance = s.balance;
    c.tree_ratio = s.tree_ratio;
    c.dynamic_domain = s.dynamic_domain ? 1 : 0;
    c.dd_alpha = s.dd_alpha;
    c.dd_radius = s.dd_radius;
    c.dd_min_radius = s.dd_min_radius;
    c.seed = s.seed;
    return c;
}

}  // namespace prrtc_detail

namespace pRRTC {

template <typename Robot>
PlannerResult<Robot> solve(typename Robot::Configuration &start, std::vector<typename Robot::Configuration> &goals,
                           ppln::collision::Environment<float> &environment, pRRTC_settings &settings) {
    constexpr int D = Robot::dimension;
    PlannerResult<Robot> res;
    prrtc_detail::SceneTables t;
    prrtc_detail::tables(environment, t);
    const prrtc_settings cs = prrtc_detail::c_settings(settings);
    std::vector<float> goal_rows;
    for (auto &g : goals) goal_rows.insert(goal_rows.end(), g.begin(), g.end());
    // Paths are a few dozen rows: solve into a per-thread buffer of 1024
    // rows, grown on demand.  A path that does not fit ends the solve with
    // PRRTC_EOVERFLOW; the planner is deterministic in (problem, settings,
    // seed), so solving again with room for the longest possible path (both
    // trees full, pRRTC.cu:552-582) returns the same plan.
    static thread_local std::vector<float> path;
    const int full = (int)std::min<long long>(2LL * settings.max_samples + 2, 1 << 20);
    int cap = std::min(full, 1024);
    prrtc_result r{};
    int rc = PRRTC_OK;
    while (true) {
        if (path.size() < (std::size_t)cap * D) path.resize((std::size_t)cap * D);
        rc = prrtc_solve(Robot::id, start.data(), goal_rows.data(), (int)goals.size(), &t.scene, &cs, &r,
                         path.data(), cap);
        if (rc != PRRTC_OK || r.status != PRRTC_EOVERFLOW || cap == full) break;
        cap = full;
    }
    if (rc != PRRTC_OK) {
        std::fprintf(stderr, "pRRTC::solve: %s\n", prrtc_last_error());
        return res;
    }
    res.solved = r.solved != 0;
    res.start_tree_size = r.start_tree_size;
    res.goal_tree_size = r.goal_tree_size;
    res.path_length = r.path_length;
    res.iters = r.iters;
    res.cost = r.cost;
    res.wall_ns = (std::size_t)r.wall_ns;
    res.kernel_ns = (std::size_t)r.kernel_ns;
    res.copy_ns = (std::size_t)r.copy_ns;
    for (int i = 0; i < r.path_nodes; i++) {
        typename Robot::Configuration c;
        std::copy_n(path.data() + (std::size_t)i * D, D, c.begin());
        res.path.push_back(c);
    }
    return res;
}

}  // namespace pRRTC
