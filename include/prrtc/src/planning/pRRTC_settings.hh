// Drop-in for pRRTC's src/planning/pRRTC_settings.hh (:3-17): the planner
// knobs with the reference's names and defaults, plus the sampler seed the
// reference fixes to 1 (pRRTC.cu:660).
#pragma once

#include <cstdint>

struct pRRTC_settings {
    int max_samples = 1000000;
    int max_iters = 1000000;
    int num_new_configs = 600;
    int granularity = 64;
    float range = 0.5;

    int balance = 1;          // 0: none; 1: proportional / tree_ratio; 2: swap when |t-o|/t < tree_ratio
    float tree_ratio = 0.5;

    bool dynamic_domain = true;
    float dd_alpha = 0.0001;
    float dd_radius = 4.0;
    float dd_min_radius = 1.0;

    uint64_t seed = 1;
};
