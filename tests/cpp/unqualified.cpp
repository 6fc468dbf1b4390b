// A harness in the style of the reference's MBM scripts: it opens only
// ppln::collision and names the robots unqualified (robots::Panda), which
// the drop-in factory header makes visible by opening ppln at file scope.
#include <cstdio>

#include "src/collision/environment.hh"
#include "src/collision/factory.hh"
#include "src/planning/Planners.hh"
#include "src/planning/pRRTC_settings.hh"

using namespace ppln::collision;

template <typename Robot>
static int run(Environment<float> &env, pRRTC_settings &settings) {
    typename Robot::Configuration start{}, goal{};
    std::vector<typename Robot::Configuration> goals{goal};
    auto res = pRRTC::solve<Robot>(start, goals, env, settings);
    return res.solved ? 0 : 1;
}

int main(int argc, char **) {
    Environment<float> env{};
    pRRTC_settings settings;
    if (argc > 5) return run<robots::Fetch>(env, settings);
    if (argc > 6) return run<robots::Baxter>(env, settings);
    if (argc > 7) return run<robots::Sphere>(env, settings);
    std::printf("robots: %d %d %d\n", robots::Panda::dimension, robots::Fetch::dimension, robots::Baxter::dimension);
    return argc > 4 ? run<robots::Panda>(env, settings) : 0;
}
