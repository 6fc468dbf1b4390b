// Drop-in for pRRTC's src/planning/Robots.hh (:10-237): robot descriptors
// with the reference's names, dimensions and joint limits (lower bound s_a,
// range s_m; a unit sample u maps to u * s_m + s_a).  The kinematics and
// collision spheres live in the HIP library; `id` selects them there.
#pragma once

#include <array>
#include <iostream>

#include "../../../prrtc.h"

namespace ppln::robots {

#define PRRTC_ROBOT_COMMON(DIM)                                                       \
    static constexpr int dimension = DIM;                                            \
    using Configuration = std::array<float, DIM>;                                    \
    static constexpr float get_s_m(int i) { return s_m[i]; }                         \
    static constexpr float get_s_a(int i) { return s_a[i]; }                         \
    static void scale_cfg(float *q) {                                                \
        for (int i = 0; i < DIM; i++) q[i] = q[i] * s_m[i] + s_a[i];                 \
    }                                                                                \
    static void print_robot_config(Configuration &cfg) {                             \
        for (int i = 0; i < DIM; i++) std::cout << cfg[i] << ' ';                    \
        std::cout << '\n';                                                           \
    }

struct Panda {
    static constexpr auto name = "panda";
    static constexpr int id = PRRTC_ROBOT_PANDA;
    static constexpr float s_m[7] = {5.9342f, 3.6652f, 5.9342f, 3.2289f, 5.9342f, 3.9096f, 5.9342f};
    static constexpr float s_a[7] = {-2.9671f, -1.8326f, -2.9671f, -3.1416f, -2.9671f, -0.0873f, -2.9671f};
    PRRTC_ROBOT_COMMON(7)
};

struct Fetch {
    static constexpr auto name = "fetch";
    static constexpr int id = PRRTC_ROBOT_FETCH;
    static constexpr float s_m[8] = {0.38615f, 3.2112f, 2.739f, 6.28318f, 4.502f, 6.28318f, 4.32f, 6.28318f};
    static constexpr float s_a[8] = {0.0f, -1.6056f, -1.221f, -3.14159f, -2.251f, -3.14159f, -2.16f, -3.14159f};
    PRRTC_ROBOT_COMMON(8)
};

struct Baxter {
    static constexpr auto name = "baxter";
    static constexpr int id = PRRTC_ROBOT_BAXTER;
    static constexpr float s_m[14] = {3.40335987756f, 3.194f, 6.10835987756f, 2.668f, 6.118f, 3.66479632679f, 6.118f,
                                      3.40335987756f, 3.194f, 6.10835987756f, 2.668f, 6.118f, 3.66479632679f, 6.118f};
    static constexpr float s_a[14] = {-1.70167993878f, -2.147f, -3.05417993878f, -0.05f, -3.059f, -1.57079632679f,
                                      -3.059f, -1.70167993878f, -2.147f, -3.05417993878f, -0.05f, -3.059f,
                                      -1.57079632679f, -3.059f};
    PRRTC_ROBOT_COMMON(14)
};

// UR5 (not in the reference): the limits of its model table, compiled from
// prrtc_amd/robots/ur5.urdf and ur5_spheres.json by the robot-model compiler.
struct UR5 {
    static constexpr auto name = "ur5";
    static constexpr int id = PRRTC_ROBOT_UR5;
    static constexpr float s_m[6] = {6.2831853f, 3.14159265f, 6.2831853f, 6.2831853f, 6.2831853f, 6.2831853f};
    static constexpr float s_a[6] = {-3.14159265f, -3.14159265f, -3.14159265f, -3.14159265f, -3.14159265f,
                                     -3.14159265f};
    PRRTC_ROBOT_COMMON(6)
};

// The 3-DoF point robot: one sphere of radius 1 at q.
struct Sphere {
    static constexpr auto name = "sphere";
    static constexpr int id = PRRTC_ROBOT_SPHERE;
    static constexpr float radius = 1.0f;
    static constexpr float s_m[3] = {10.0f, 10.0f, 10.0f};
    static constexpr float s_a[3] = {-5.0f, -5.0f, -5.0f};
    PRRTC_ROBOT_COMMON(3)
};

#undef PRRTC_ROBOT_COMMON

}  // namespace ppln::robots
