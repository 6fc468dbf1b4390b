// Drop-in for pRRTC's src/collision/factory.hh without Eigen: obstacle
// constructors from centre / Euler XYZ angles / extents (factory.hh:14-429).
//
//   cuboid::flat / array       box from centre, Euler XYZ (rho, theta, phi)
//                              and half-extents; R = Rz(phi) Ry(theta) Rx(rho),
//                              box axes = columns of R (factory.hh:28-63)
//   cylinder::endpoints::*     capsule between two points, rdv = 1/|v|^2
//                              (factory.hh:118-129)
//   cylinder::center::*        capsule of length L about a centre, end points
//                              +-L/2 along the rotated z axis (factory.hh:165-185)
//   capsule::*                 the same constructors under the capsule name
//   sphere::flat / array       (factory.hh:351-355)
//
// The Eigen-typed overloads of the reference are not provided.
#pragma once

#include <array>
#include <cmath>

#include "shapes.hh"

// The reference header opens ppln at file scope (factory.hh:12); harnesses
// written against it name robots::Panda and the like unqualified.
using namespace ppln;

namespace ppln::collision::factory {

using ConstArrayRef = const std::array<float, 3> &;

namespace detail {
// R = Rz(phi) * Ry(theta) * Rx(rho), row-major
inline std::array<float, 9> euler_xyz(float rho, float theta, float phi) {
    const float cr = std::cos(rho), sr = std::sin(rho);
    const float ct = std::cos(theta), st = std::sin(theta);
    const float cp = std::cos(phi), sp = std::sin(phi);
    return {cp * ct, cp * st * sr - sp * cr, cp * st * cr + sp * sr,
            sp * ct, sp * st * sr + cp * cr, sp * st * cr - cp * sr,
            -st,     ct * sr,                ct * cr};
}
}  // namespace detail

namespace cuboid {
inline Cuboid<float> flat(float cx, float cy, float cz, float rho, float theta, float phi, float hx, float hy,
                          float hz) noexcept {
    const auto R = detail::euler_xyz(rho, theta, phi);
    return Cuboid<float>(cx, cy, cz, R[0], R[3], R[6], R[1], R[4], R[7], R[2], R[5], R[8], hx, hy, hz);
}
inline Cuboid<float> array(ConstArrayRef center, ConstArrayRef euler_xyz, ConstArrayRef half_extents) noexcept {
    return flat(center[0], center[1], center[2], euler_xyz[0], euler_xyz[1], euler_xyz[2], half_extents[0],
                half_extents[1], half_extents[2]);
}
}  // namespace cuboid

namespace cylinder {
namespace endpoints {
inline Cylinder<float> flat(float x1, float y1, float z1, float x2, float y2, float z2, float radius) noexcept {
    const float xv = x2 - x1, yv = y2 - y1, zv = z2 - z1;
    const float len2 = xv * xv + yv * yv + zv * zv;
    return Cylinder<float>(x1, y1, z1, xv, yv, zv, radius, static_cast<float>(1.0 / len2));
}
inline Cylinder<float> array(ConstArrayRef a, ConstArrayRef b, float radius) noexcept {
    return flat(a[0], a[1], a[2], b[0], b[1], b[2], radius);
}
}  // namespace endpoints

namespace center {
inline Cylinder<float> flat(float cx, float cy, float cz, float rho, float theta, float phi, float radius,
                            float length) noexcept {
    const auto R = detail::euler_xyz(rho, theta, phi);
    const float h = length / 2;
    const float zx = R[2] * h, zy = R[5] * h, zz = R[8] * h;   // rotated +z, scaled
    return endpoints::flat(cx + zx, cy + zy, cz + zz, cx - zx, cy - zy, cz - zz, radius);
}
inline Cylinder<float> array(ConstArrayRef center, ConstArrayRef euler_xyz, float radius, float length) noexcept {
    return flat(center[0], center[1], center[2], euler_xyz[0], euler_xyz[1], euler_xyz[2], radius, length);
}
}  // namespace center
}  // namespace cylinder

namespace capsule {
namespace endpoints = cylinder::endpoints;
namespace center = cylinder::center;
}  // namespace capsule

namespace sphere {
inline Sphere<float> flat(float x, float y, float z, float radius) noexcept { return Sphere<float>(x, y, z, radius); }
inline Sphere<float> array(ConstArrayRef c, float radius) noexcept { return flat(c[0], c[1], c[2], radius); }
}  // namespace sphere

}  // namespace ppln::collision::factory
