// Nearest-node search over SoA trees on gfx950: the planner's scan
// (nearest<D>, planner_device.h) exposed for measurement and parity.
//
// Replaces the per-block tree scans of the reference planner
// (src/planning/pRRTC.cu:384-423 extend, 478-514 connect; distance at
// src/planning/utils.cuh:89-109): there each of the 128 threads of a block
// strides the AoS node array and a shared-memory reduction picks the minimum.
// Here one workgroup answers one query: its waves stride the tree's SoA
// columns in 1024-node chunks of float4 loads and reduce by lane shuffles
// and LDS, ties going to the smaller index.
#include <hip/hip_runtime.h>

#include "host_common.h"
#include "planner_device.h"

namespace prrtc {

// Query k = t * qpt + j asks tree t (`trees + t * D * cap`) for the node
// nearest to queries[k]; idx[k] and dist2[k] receive the answer.
template <int D>
__global__ __launch_bounds__(kMaxTaskWaves * kWave) void k_nearest(const float *__restrict__ trees, int cap, int n,
                                                                    const float *__restrict__ queries, int qpt,
                                                                    int n_queries, int *__restrict__ idx,
                                                                    float *__restrict__ dist2) {
    __shared__ float red[2 * kMaxTaskWaves];
    // Workgroups reach the XCDs round-robin (workgroup b runs on XCD b % 8)
    // and each XCD has its own L2.  The queries are split into 8 contiguous
    // ranges, one per XCD, so the queries of one tree (consecutive k) share
    // that tree's lines in one L2 instead of fetching them into all eight.
    int lo = 0, hi = n_queries, first = blockIdx.x, stride = gridDim.x;
    if ((int)gridDim.x >= kXcds) {
        const int x = blockIdx.x % kXcds;
        lo = (int)((long long)n_queries * x / kXcds);
        hi = (int)((long long)n_queries * (x + 1) / kXcds);
        first = blockIdx.x / kXcds;
        stride = ((int)gridDim.x - x + kXcds - 1) / kXcds;   // workgroups on this XCD
    }
    for (int k = lo + first; k < hi; k += stride) {
        const float *tree = trees + (size_t)(k / qpt) * D * cap;
        float q[D];
#pragma unroll
        for (int d = 0; d < D; d++) q[d] = queries[(size_t)k * D + d];
        float best;
        const int i = nearest<D>(tree, cap, n, q, best, red, team_workgroup());
        if (threadIdx.x == 0) {
            idx[k] = i;
            dist2[k] = best;
        }
    }
}

template <int D>
static void launch_nearest(const float *trees, int cap, int n, const float *queries, int qpt, int n_queries,
                           int waves, int *idx, float *dist2, hipStream_t s) {
    const int grid = std::max(1, std::min(n_queries, num_cus() * 64));
    hipLaunchKernelGGL(k_nearest<D>, dim3(grid), dim3(waves * kWave), 0, s, trees, cap, n, queries, qpt, n_queries,
                       idx, dist2);
    PRRTC_HIP(hipGetLastError());
}

}  // namespace prrtc

using namespace prrtc;

extern "C" int prrtc_nearest_async(int dof, const float *d_trees, int cap, int n_nodes, int n_trees,
                                   const float *d_queries, int queries_per_tree, int waves, int *d_idx,
                                   float *d_dist2, void *stream) {
    return guarded("prrtc_nearest_async", [&]() -> int {
        if ((dof != 3 && dof != 7 && dof != 8 && dof != 14) || n_trees < 0 || queries_per_tree < 0 ||
            n_nodes < 1 || cap < n_nodes || cap % kScanChunk != 0 || waves < 1 || waves > kMaxTaskWaves ||
            (n_trees > 0 && queries_per_tree > 0 && (!d_trees || !d_queries || !d_idx || !d_dist2))) {
            set_error("prrtc_nearest_async: invalid argument (dof 3/7/8/14, 1 <= n_nodes <= cap, cap a multiple "
                      "of %d, 1 <= waves <= %d)", kScanChunk, kMaxTaskWaves);
            return PRRTC_EINVAL;
        }
        const long long nq = (long long)n_trees * queries_per_tree;
        if (nq == 0) return PRRTC_OK;
        if (nq > INT32_MAX) {
            set_error("prrtc_nearest_async: too many queries");
            return PRRTC_EINVAL;
        }
        const hipStream_t s = static_cast<hipStream_t>(stream);
        switch (dof) {
            case 3: launch_nearest<3>(d_trees, cap, n_nodes, d_queries, queries_per_tree, (int)nq, waves, d_idx,
                                      d_dist2, s); break;
            case 7: launch_nearest<7>(d_trees, cap, n_nodes, d_queries, queries_per_tree, (int)nq, waves, d_idx,
                                      d_dist2, s); break;
            case 8: launch_nearest<8>(d_trees, cap, n_nodes, d_queries, queries_per_tree, (int)nq, waves, d_idx,
                                      d_dist2, s); break;
            case 14: launch_nearest<14>(d_trees, cap, n_nodes, d_queries, queries_per_tree, (int)nq, waves, d_idx,
                                        d_dist2, s); break;
        }
        return PRRTC_OK;
    });
}
