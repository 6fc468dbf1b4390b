// A harness written against the reference's headers (the include paths and
// calls of scripts/evaluate_mbm.cpp) compiled against the drop-in headers and
// libprrtc.so.  Solves one Panda problem among spheres, a capsule and a box;
// prints "solved cost path_length nodes" or exits non-zero.
#include <cstdio>

#include "src/collision/environment.hh"
#include "src/collision/factory.hh"
#include "src/planning/Planners.hh"
#include "src/planning/pRRTC_settings.hh"

using namespace ppln::collision;

int main() {
    Environment<float> env{};
    env.spheres = new Sphere<float>[2]{Sphere<float>(0.55f, 0.1f, 0.5f, 0.08f), Sphere<float>(-0.3f, 0.5f, 0.9f, 0.12f)};
    env.num_spheres = 2;
    env.capsules = new Capsule<float>[1]{factory::cylinder::center::array({0.4f, -0.5f, 0.6f}, {0.f, 0.f, 0.f}, 0.05f, 0.4f)};
    env.num_capsules = 1;
    env.cuboids = new Cuboid<float>[1]{factory::cuboid::array({-0.5f, -0.4f, 0.3f}, {0.3f, 0.1f, 0.7f}, {0.1f, 0.05f, 0.2f})};
    env.num_cuboids = 1;
    using Robot = ppln::robots::Panda;
    Robot::Configuration start = {0.f, -0.785f, 0.f, -2.356f, 0.f, 1.571f, 0.785f};
    std::vector<Robot::Configuration> goals = {{2.35f, 1.f, 0.f, -0.8f, 0.f, 2.5f, 0.785f}};
    pRRTC_settings settings;
    settings.num_new_configs = 64;
    settings.granularity = 32;
    settings.max_iters = 2000;
    settings.max_samples = 100000;
    auto res = pRRTC::solve<Robot>(start, goals, env, settings);
    if (!res.solved) return 2;
    if (res.path.front() != goals[0] || res.path.back() != start) return 3;
    std::printf("%d %f %d %zu\n", (int)res.solved, res.cost, res.path_length, res.path.size());
    return 0;
}
