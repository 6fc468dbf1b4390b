/* Sizes and field offsets of the C ABI structs, for the ctypes mirror test. */
#include <stddef.h>
#include <stdio.h>

#include "prrtc.h"

#define S(T) printf(#T " %zu\n", sizeof(T))
#define O(T, F) printf(#T "." #F " %zu\n", offsetof(T, F))

int main(void) {
    S(prrtc_scene); S(prrtc_settings); S(prrtc_result); S(prrtc_problem); S(prrtc_batch_stats);
    S(prrtc_tuning); S(prrtc_server_config); S(prrtc_server_stats); S(prrtc_request_timing);
    O(prrtc_result, edge_checks); O(prrtc_result, pose_checks); O(prrtc_batch_stats, pose_checks);
    O(prrtc_server_config, deadline_ms); O(prrtc_server_stats, timed_out); O(prrtc_settings, seed);
    return 0;
}
