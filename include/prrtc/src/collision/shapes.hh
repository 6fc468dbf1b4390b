// Drop-in for pRRTC's src/collision/shapes.hh (shapes.hh:13-245): the
// obstacle types with the same names, fields and constructors.  Each shape
// carries a name and its distance from the origin (`min_distance`).
#pragma once

#include <cmath>
#include <string>

#include "math.hh"

namespace ppln::collision {

template <typename DataT>
struct Shape {
    std::string name;
    DataT min_distance{};
};

// Box centred at (x, y, z) with axes axis_1..3 (unit, orthogonal) and
// half-extents axis_1_r..axis_3_r along them.
template <typename DataT>
struct Cuboid : public Shape<DataT> {
    DataT x{}, y{}, z{};
    DataT axis_1_x{}, axis_1_y{}, axis_1_z{};
    DataT axis_2_x{}, axis_2_y{}, axis_2_z{};
    DataT axis_3_x{}, axis_3_y{}, axis_3_z{};
    DataT axis_1_r{}, axis_2_r{}, axis_3_r{};

    Cuboid() = default;
    Cuboid(DataT x_, DataT y_, DataT z_, DataT a1x, DataT a1y, DataT a1z, DataT a2x, DataT a2y, DataT a2z, DataT a3x,
           DataT a3y, DataT a3z, DataT r1, DataT r2, DataT r3)
        : x(x_), y(y_), z(z_), axis_1_x(a1x), axis_1_y(a1y), axis_1_z(a1z), axis_2_x(a2x), axis_2_y(a2y),
          axis_2_z(a2z), axis_3_x(a3x), axis_3_y(a3y), axis_3_z(a3z), axis_1_r(r1), axis_2_r(r2), axis_3_r(r3) {
        // closest point of the box to the origin, expressed in the box frame
        const DataT p1 = clamp(-dot_3(x, y, z, axis_1_x, axis_1_y, axis_1_z), -axis_1_r, axis_1_r);
        const DataT p2 = clamp(-dot_3(x, y, z, axis_2_x, axis_2_y, axis_2_z), -axis_2_r, axis_2_r);
        const DataT p3 = clamp(-dot_3(x, y, z, axis_3_x, axis_3_y, axis_3_z), -axis_3_r, axis_3_r);
        const DataT cx = x + axis_1_x * p1 + axis_2_x * p2 + axis_3_x * p3;
        const DataT cy = y + axis_1_y * p1 + axis_2_y * p2 + axis_3_y * p3;
        const DataT cz = z + axis_1_z * p1 + axis_2_z * p2 + axis_3_z * p3;
        this->min_distance = std::sqrt(cx * cx + cy * cy + cz * cz);
    }
};

// Cylinder / capsule: first end point (x1, y1, z1), vector to the second end
// point (xv, yv, zv), radius r, and rdv = 1 / |v|^2.
template <typename DataT>
struct Cylinder : public Shape<DataT> {
    DataT x1{}, y1{}, z1{};
    DataT xv{}, yv{}, zv{};
    DataT r{};
    DataT rdv{};

    Cylinder() = default;
    Cylinder(DataT x1_, DataT y1_, DataT z1_, DataT xv_, DataT yv_, DataT zv_, DataT r_, DataT rdv_)
        : x1(x1_), y1(y1_), z1(z1_), xv(xv_), yv(yv_), zv(zv_), r(r_), rdv(rdv_) {
        // distance from the origin to the capsule surface
        const DataT t = clamp(-dot_3(x1, y1, z1, xv, yv, zv) * rdv, DataT(0), DataT(1));
        const DataT px = x1 + xv * t, py = y1 + yv * t, pz = z1 + zv * t;
        const DataT axis = std::sqrt(px * px + py * py + pz * pz);
        this->min_distance = std::max(DataT(0), axis - r);
    }
    DataT x2() const { return x1 + xv; }
    DataT y2() const { return y1 + yv; }
    DataT z2() const { return z1 + zv; }
};

template <typename DataT>
using Capsule = Cylinder<DataT>;

template <typename DataT>
struct Sphere : public Shape<DataT> {
    DataT x{}, y{}, z{}, r{};

    Sphere() = default;
    Sphere(DataT x_, DataT y_, DataT z_, DataT r_) : x(x_), y(y_), z(z_), r(r_) {
        this->min_distance = std::sqrt(x * x + y * y + z * z) - r;
    }
};

}  // namespace ppln::collision
