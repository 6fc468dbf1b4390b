// Host side of the planner shared by the batched solver (planner.hip) and
// the serving loop (server.hip): request validation, the per-problem device
// block layout, and launch wrappers that dispatch on the robot id.
#pragma once

#include "host_common.h"
#include "planner_device.h"

namespace prrtc {

// One planning problem as handed to the C ABI (pointers stay the caller's).
struct HostProblem {
    int robot;
    const float *start;
    const float *goals;
    int n_goals;
    const prrtc_scene *scene;
    const prrtc_settings *st;
    int index;                  // position in the caller's arrays
};

// Device block of one problem: everything its PlanDesc points to.  The
// packed scene and the init rows (start, goals) lead the block so one copy
// uploads both.
struct BlockLayout {
    size_t scene = 0, init = 0, nodes = 0, parent = 0, radii = 0, halton = 0, tprev = 0, meta = 0, stage = 0;
    size_t upload_bytes = 0;    // scene + init, contiguous at the block start
    size_t bytes = 0;           // whole block (256-byte multiple)
    int cap = 0, max_ext = 0, scene_floats = 0;
};

// Internal robot id of a launch whose problems have different robots (one
// persistent grid for all of them; never seen by callers of the C ABI).
constexpr int kRobotMixed = kNumRobots;
// Floats per configuration of a launch: the robot's DOF, or the largest DOF
// for kRobotMixed (the layout of its task areas and path rows).
int plan_dof(int robot);

// Rejects what the planner cannot run (unknown robot, NaN, oversized scene,
// bad settings, chains over kMaxExt); sets the thread's error message.
int check_problem(const HostProblem &hp, const char *fn);
int robot_limits(int robot, float *lo, float *span);
int max_extensions(const HostProblem &hp);
BlockLayout block_layout(const HostProblem &hp);
// Writes the packed scene and init rows to `host` (layout.upload_bytes).
void block_upload_image(const HostProblem &hp, const BlockLayout &L, char *host);
// Fills a PlanDesc whose buffers live in the device block at `dev`.
void fill_desc(PlanDesc &d, const HostProblem &hp, const BlockLayout &L, char *dev);
// PlanControl::wide_slots of a launch on this device
int plan_wide_slots();

// Launch shapes of a robot's persistent planner grids: the value set by
// prrtc_set_tuning, else an environment variable, else the default.
//   plan_task_waves    waves per task of a single solve, or of a server
//                      slice with few resident problems (PRRTC_TASK_WAVES)
//   plan_batch_waves   waves per task of a batch, or of a server slice with
//                      many resident problems (PRRTC_BATCH_WAVES)
//   plan_waves_per_cu  waves per CU of a grid (PRRTC_WAVES_PER_CU)
//   plan_grid          task workgroups of a grid with `waves` per task
constexpr int kMaxWavesPerCu = 32;
constexpr int kDefaultTaskWaves = 4;
constexpr int kDefaultBatchWaves = 4;
constexpr int kDefaultWavesPerCu = 8;
// planner waves a CU holds at once: the real-robot k_plan kernels need
// 256 VGPRs plus AGPRs, one wave per SIMD
constexpr int kResidentPlanWavesPerCu = 4;
int plan_task_waves(int robot);
int plan_batch_waves(int robot);
int plan_waves_per_cu(int robot);
int plan_grid(int robot, int waves, long long slots);
void run_round_graph(int robot, const PlanDesc *d_desc, PlanState *d_state, PlanControl *d_ctrl, float *d_paths,
                     int P, int max_nnc, int waves, int lds_scene_f4, int max_ext, const unsigned long long *d_run,
                     unsigned long long *h_run, int words, long long max_rounds, hipStream_t s);
// Resident problems up to which a server slice runs the single-solve shape
// (waves_per_task) instead of the batch shape: as many as waves_per_cu x
// CUs one-wave tasks hold with every slot in flight.
int plan_tail_problems(int robot, int max_nnc);
// Running problems above which a workgroup claims one slot per wave
// (PlanControl::crowd_above; PRRTC_CROWD_ABOVE, < 0 turns it off).
int plan_crowd_above(int robot, int max_nnc);

// Tree roots and Halton states of the problems listed in d_list (table
// entries), or of entries 0..n-1 when d_list is null.
void launch_plan_init(int robot, const PlanDesc *d_desc, PlanState *d_state, const int *d_list, int n,
                      hipStream_t s);
// One persistent launch over the problem table (runs until every running
// problem stops, or until the control block's time slice ends).
void launch_plan_slice(int robot, const PlanDesc *d_desc, PlanState *d_state, PlanControl *d_ctrl, float *d_paths,
                       int grid, int waves, int lds_scene_f4, int max_ext, hipStream_t s);

// Rows of one solution: goal -> start (pRRTC.cu:715-746) from the two traces
// the commit wrote at state.path_off, `rs` floats per row in `rows` (the
// launch's PlanControl::path_stride); `out` gets the robot's DOF per row.
// Returns the row count, or -1 if the path does not fit path_cap rows.
int assemble_path(const HostProblem &hp, const PlanState &ps, const float *rows, int rs, int path_cap, float *out);

}  // namespace prrtc
