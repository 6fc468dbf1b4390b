/*
 * prrtc -- C ABI of the MI355X (gfx950) pRRTC planning hot path.
 *
 * Plain C types only: float/int pointers and sizes.  Every function returns
 * PRRTC_OK (0) or a negative PRRTC_E* code and never exits the process;
 * prrtc_last_error() gives the message of the calling thread's last failure.
 *
 * Each entry point names the reference interface it replaces
 * (pranavjad/pRRTC, paths relative to the repository root).
 */
#ifndef PRRTC_H
#define PRRTC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRRTC_OK 0
#define PRRTC_EINVAL -1      /* bad argument (robot id, size, NULL pointer, NaN) */
#define PRRTC_EDEVICE -2     /* HIP runtime error */
#define PRRTC_ENOMEM -3      /* device allocation / admission budget exceeded */
#define PRRTC_EOVERFLOW -4   /* a caller buffer (path, result) is too small */
#define PRRTC_EBUSY -5       /* server queue full */
#define PRRTC_EAGAIN -6      /* server result not ready yet */
#define PRRTC_ETIMEDOUT -7   /* server request passed its deadline (result status) */
#define PRRTC_ECANCELED -8   /* server request cancelled (result status) */

/* Robot ids: the reference's explicit instantiations (src/planning/pRRTC.cu:768-771). */
#define PRRTC_ROBOT_SPHERE 0 /* 3-DoF point robot, radius 1 (src/planning/Robots.hh:190-237) */
#define PRRTC_ROBOT_PANDA 1  /* 7 DoF  (Robots.hh:10-64)   */
#define PRRTC_ROBOT_FETCH 2  /* 8 DoF  (Robots.hh:66-120)  */
#define PRRTC_ROBOT_BAXTER 3 /* 14 DoF (Robots.hh:122-188) */
#define PRRTC_ROBOT_UR5 4    /* 6 DoF: not in the reference; compiled from prrtc_amd/robots/ur5.urdf +
                                ur5_spheres.json by the robot-model compiler */

/* Obstacle tables, row-major float32.  The five tables the reference's
 * checker reads (src/planning/utils.cuh:211-242; shapes at
 * src/collision/shapes.hh:13-245, environment at src/collision/environment.hh:12-45):
 *   spheres    [n][4]  x y z r
 *   capsules   [n][8]  x1 y1 z1 xv yv zv r rdv    (rdv = 1/|v|^2; "cylinders" are capsules)
 *   z_capsules [n][8]  same, v along z
 *   cuboids    [n][15] x y z a1(xyz) a2(xyz) a3(xyz) r1 r2 r3
 *   z_cuboids  [n][15] same, a3 = +z                                      */
typedef struct prrtc_scene {
    const float *spheres;    int32_t n_spheres;
    const float *capsules;   int32_t n_capsules;
    const float *z_capsules; int32_t n_z_capsules;
    const float *cuboids;    int32_t n_cuboids;
    const float *z_cuboids;  int32_t n_z_cuboids;
} prrtc_scene;

/* pRRTC_settings (src/planning/pRRTC_settings.hh:3-17), same fields and
 * meaning, fixed-width, plus the sampler seed (the reference fixes it to 1,
 * src/planning/pRRTC.cu:660). */
typedef struct prrtc_settings {
    int32_t max_samples;
    int32_t max_iters;
    int32_t num_new_configs;
    int32_t granularity;
    float range;
    int32_t balance;
    float tree_ratio;
    int32_t dynamic_domain;
    float dd_alpha;
    float dd_radius;
    float dd_min_radius;
    uint64_t seed;
} prrtc_settings;

/* PlannerResult (src/planning/Planners.hh:11-23) without the path vector;
 * the path goes to a caller buffer, goal -> start, as rows of dof floats. */
typedef struct prrtc_result {
    int32_t solved;
    int32_t start_tree_size;
    int32_t goal_tree_size;
    int32_t path_length;      /* interior nodes: path_nodes - 2 (pRRTC.cu:743) */
    int32_t iters;
    float cost;
    int32_t path_nodes;       /* rows written to the path buffer */
    int32_t status;           /* PRRTC_OK or why the problem ended unsolved */
    int64_t wall_ns;          /* API call -> result (pRRTC.cu:615, 763) */
    int64_t kernel_ns;        /* device time of the solve (pRRTC.cu:696-706) */
    int64_t copy_ns;          /* host<->device copies (pRRTC.cu:685-694, 709-713) */
    int64_t edge_checks;      /* edges whose verdict was decided */
    int64_t pose_checks;      /* poses put on lanes for those edges (one pose per lane and pass;
                                 like the times, it depends on the launch shape) */
} prrtc_result;

/* Fill `s` with the reference's defaults (pRRTC_settings.hh:4-16), seed 1. */
void prrtc_settings_default(prrtc_settings *s);

int prrtc_version(void);
const char *prrtc_last_error(void);
int prrtc_robot_dof(int robot);
int prrtc_robot_num_spheres(int robot);

/* ---- scenes ------------------------------------------------------------ */
/* Packed device form of a scene (float4-aligned tables, counts in a header).
 * Each cuboid also carries a cull radius computed here (its padded
 * half-diagonal), so device buffers for the *_async entry points must be
 * written by prrtc_scene_pack, not assembled by hand. */
size_t prrtc_scene_packed_floats(const prrtc_scene *scene);
int prrtc_scene_pack(const prrtc_scene *scene, float *out, size_t cap_floats);

/* ---- FK + collision (replaces ppln::collision::fkcc<Robot>, utils.cuh:326-33131) */
/* Sphere centres of every collision sphere, [n][num_spheres][3] (debug/parity). */
int prrtc_fk_spheres(int robot, const float *poses, int n, float *out_centres);
/* Per-pose verdict, 1 = collision-free. */
int prrtc_check_poses(int robot, const prrtc_scene *scene, const float *poses, int n, uint8_t *out_free);

/* ---- edge validation (replaces the per-edge block of pRRTC.cu:430-445) ----
 * edges [n][2][dof]; edge (a,b) is free iff poses a + (k+1)(b-a)/g,
 * k = 0..g-1, are all free.  Host buffers, synchronous. */
int prrtc_validate_edges(int robot, const prrtc_scene *scene, const float *edges, int n, int granularity,
                         uint8_t *out_free);
/* Same on device buffers, asynchronous on `stream` (a hipStream_t, NULL =
 * default stream).  d_scene is a packed scene (prrtc_scene_pack) of
 * scene_floats floats in device memory. */
int prrtc_validate_edges_async(int robot, const float *d_scene, size_t scene_floats, const float *d_edges, int n,
                               int granularity, uint8_t *d_out_free, void *stream);

/* ---- nearest-node search (replaces the tree scans of pRRTC.cu:384-423 and
 *      478-514, distance utils.cuh:89-109), device buffers, asynchronous ----
 * n_trees SoA trees, tree t at d_trees + t * dof * cap: coordinate k of node
 * i at [k * cap + i], nodes [0, n_nodes) valid, cap a multiple of 1024.
 * Query j of tree t is d_queries[(t * queries_per_tree + j) * dof ..]; its
 * answer (smallest squared L2 distance, ties to the smaller index) goes to
 * d_idx / d_dist2 at t * queries_per_tree + j.  waves (1..4) wavefronts
 * share one query.  dof is 3, 7, 8 or 14. */
int prrtc_nearest_async(int dof, const float *d_trees, int cap, int n_nodes, int n_trees, const float *d_queries,
                        int queries_per_tree, int waves, int *d_idx, float *d_dist2, void *stream);

/* ---- planner (replaces pRRTC::solve<Robot>, src/planning/Planners.hh:75-78,
 *      defined at src/planning/pRRTC.cu:607-766) ----
 * path: path_cap rows of dof floats, written goal -> start. */
int prrtc_solve(int robot, const float *start, const float *goals, int n_goals, const prrtc_scene *scene,
                const prrtc_settings *settings, prrtc_result *result, float *path, int path_cap);

/* prrtc_solve without allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * All device state of the solve (trees, slot states, staged chains, path
 * rows) lives in one static device array of PRRTC_NOALLOC_SCRATCH_BYTES per
 * GPU, held until the process ends; the call runs on the null stream and
 * returns when done, with the result prrtc_solve returns, bit for bit.  A
 * problem whose state does not fit (mostly max_samples and path_cap) fails
 * with PRRTC_ENOMEM and a message naming the sizes.  Calls on one GPU share
 * the array: calls that overlap there (several host threads) race, so make
 * at most one at a time per GPU. */
#define PRRTC_NOALLOC_SCRATCH_BYTES (64u << 20)
int prrtc_solve_noalloc(int robot, const float *start, const float *goals, int n_goals, const prrtc_scene *scene,
                        const prrtc_settings *settings, prrtc_result *result, float *path, int path_cap);

/* Many independent problems in one persistent launch (no reference
 * counterpart; the reference solves one problem per launch). */
typedef struct prrtc_problem {
    int32_t robot;
    const float *start;       /* dof floats */
    const float *goals;       /* n_goals * dof floats */
    int32_t n_goals;
    prrtc_scene scene;
} prrtc_problem;

typedef struct prrtc_batch_stats {
    int64_t wall_ns;          /* whole call */
    int64_t kernel_ns;        /* device time of the persistent launch */
    int64_t edge_checks;      /* over all problems */
    int32_t solved;           /* problems solved */
    int32_t rounds;           /* planner rounds executed (max over problems) */
    int64_t pose_checks;      /* over all problems */
} prrtc_batch_stats;

/* settings[i] applies to problem i (settings may also point to one struct
 * when n_settings == 1).  paths: one block of path_cap * 14 floats per
 * problem; block i holds results[i].path_nodes rows of that problem's dof
 * floats, packed, goal -> start.  The problems may have different robots:
 * one persistent grid serves them all, and each result equals that
 * problem's prrtc_solve. */
int prrtc_solve_batch(int n_problems, const prrtc_problem *problems, const prrtc_settings *settings,
                      int n_settings, prrtc_result *results, float *paths, int path_cap,
                      prrtc_batch_stats *stats);

/* The same batch without a persistent kernel (no reference counterpart;
 * the non-persistent comparator of prrtc_solve_batch): every round is a grid
 * of all running problems' tasks followed by a commit grid, eight rounds per
 * captured graph, replayed until no problem runs.  Same arguments, same
 * results; kernel_ns covers the replays. */
int prrtc_solve_batch_rounds(int n_problems, const prrtc_problem *problems, const prrtc_settings *settings,
                             int n_settings, prrtc_result *results, float *paths, int path_cap,
                             prrtc_batch_stats *stats);

/* Launch shape of the persistent planner per robot (no reference
 * counterpart: the reference launches num_new_configs blocks of 128 threads
 * per problem, src/planning/pRRTC.cu:700-702).  0 = default for each field.
 *   waves_per_cu          waves per CU of a persistent grid, 1..32
 *                         (environment PRRTC_WAVES_PER_CU, else 8)
 *   waves_per_task        waves (64 lanes each) sharing one planner task of a
 *                         single solve, or of a server slice with few
 *                         resident problems: they share the nearest-node
 *                         scans and split the obstacle tables of every
 *                         validation pass (1..4; environment
 *                         PRRTC_TASK_WAVES, else 4)
 *   batch_waves_per_task  the same for prrtc_solve_batch and for server
 *                         slices with many resident problems (1..4;
 *                         PRRTC_BATCH_WAVES, else 4)
 *   crowd_above           with more running problems than this, a batch or
 *                         server workgroup claims one slot per wave and each
 *                         wave runs its own task with the whole scene
 *                         (throughput); with fewer its waves share one slot
 *                         (latency).  < 0: never; 0: PRRTC_CROWD_ABOVE, else
 *                         CUs x 4 / num_new_configs (the planner holds one
 *                         wave per SIMD).  prrtc_get_tuning returns the
 *                         setting (0 = that default)
 * robot = -1 sets every robot.  Applies to later prrtc_solve,
 * prrtc_solve_batch and server slices; results do not depend on it. */
typedef struct prrtc_tuning {
    int32_t waves_per_cu;
    int32_t waves_per_task;
    int32_t batch_waves_per_task;
    int32_t crowd_above;
} prrtc_tuning;

int prrtc_set_tuning(int robot, const prrtc_tuning *t);
int prrtc_get_tuning(int robot, prrtc_tuning *t);

/* ---- serving (no reference counterpart: the reference serves one blocking
 *      solve per process at a time, src/planning/pRRTC.cu:607-766) ----
 * A server owns a device-memory budget and a worker thread.  Requests queue
 * in arrival order and are admitted while their device footprint fits the
 * budget; admitted problems run in time-sliced persistent launches (one
 * stream per robot) and keep their state in device memory between slices.
 * A request's result (path, iters, tree sizes, cost) is the one prrtc_solve
 * returns for the same inputs, whatever else is resident.  All server calls
 * are thread-safe. */
typedef struct prrtc_server prrtc_server;

typedef struct prrtc_server_config {
    uint64_t budget_bytes;    /* device memory for resident problems (0: 2 GiB) */
    int32_t max_resident;     /* resident problems per robot (0: 1024) */
    int32_t max_pending;      /* queued requests before submit returns EBUSY (0: 65536) */
    int32_t slice_us;         /* length of one persistent launch (0: 1000) */
    int32_t path_cap;         /* rows kept per result path (0: 1024) */
    /* a request's slots are claimed ahead of older requests' for its first
     * young_rounds planner rounds (0: 64; < 0: never, plain round-robin) */
    int32_t young_rounds;
    /* a request still unsolved this long after its submit ends unsolved
     * with result status PRRTC_ETIMEDOUT (0: no deadline).  A queued request
     * is dropped without running; a resident one stops at the end of its
     * current planner round (its remaining tasks of that round are skipped) */
    int32_t deadline_ms;
} prrtc_server_config;

typedef struct prrtc_server_stats {
    int64_t submitted;
    int64_t completed;        /* results produced (polled or not) */
    int64_t solved;
    int32_t pending;          /* queued, not yet admitted */
    int32_t resident;         /* admitted, not finished */
    uint64_t resident_bytes;  /* budget in use */
    uint64_t peak_resident_bytes;
    int64_t slices;           /* persistent launches */
    int64_t kernel_ns;        /* device time over all slices */
    int64_t edge_checks;
    int64_t timed_out;        /* requests ended by the deadline */
    int64_t cancelled;        /* requests ended by prrtc_server_cancel */
} prrtc_server_stats;

typedef struct prrtc_request_timing {
    int64_t queued_ns;        /* submit -> admission */
    int64_t resident_ns;      /* admission -> result */
    int32_t slices;           /* persistent launches the request ran in */
} prrtc_request_timing;

void prrtc_server_config_default(prrtc_server_config *cfg);
int prrtc_server_create(const prrtc_server_config *cfg, prrtc_server **out);
/* Copies the problem (start, goals, scene) and settings; *ticket names the
 * request (> 0).  EBUSY: queue full; ENOMEM: footprint above the budget. */
int prrtc_server_submit(prrtc_server *srv, const prrtc_problem *problem, const prrtc_settings *settings,
                        uint64_t *ticket);
/* Waits up to wait_ms (< 0: no limit) for the ticket's result.  PRRTC_OK:
 * result, path (rows goal -> start, at most path_cap) and timing (may be
 * NULL) written and the ticket released; EAGAIN: not finished yet. */
int prrtc_server_poll(prrtc_server *srv, uint64_t ticket, int32_t wait_ms, prrtc_result *result, float *path,
                      int32_t path_cap, prrtc_request_timing *timing);
int prrtc_server_get_stats(prrtc_server *srv, prrtc_server_stats *out);
/* Ends a request unsolved with result status PRRTC_ECANCELED: at once if it
 * is still queued, else at the end of its current planner round.  A request
 * that has already finished keeps its result (PRRTC_OK, nothing changes);
 * EINVAL for a ticket the server does not hold.  Poll the ticket as usual. */
int prrtc_server_cancel(prrtc_server *srv, uint64_t ticket);
/* Stops the worker after its current slice and frees the server; queued,
 * resident and unpolled requests are dropped. */
int prrtc_server_destroy(prrtc_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* PRRTC_H */
