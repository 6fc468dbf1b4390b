// Planner entry points (filled in by the round-synchronous planner).
#include "host_common.h"

using namespace prrtc;

extern "C" {

int prrtc_solve(int robot, const float *start, const float *goals, int n_goals, const prrtc_scene *scene,
                const prrtc_settings *settings, prrtc_result *result, float *path, int path_cap) {
    set_error("prrtc_solve: planner not available in this build");
    return PRRTC_EINVAL;
}

int prrtc_solve_batch(int n_problems, const prrtc_problem *problems, const prrtc_settings *settings,
                      int n_settings, prrtc_result *results, float *paths, int path_cap,
                      prrtc_batch_stats *stats) {
    set_error("prrtc_solve_batch: planner not available in this build");
    return PRRTC_EINVAL;
}

}  // extern "C"
