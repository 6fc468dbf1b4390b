// Drop-in for pRRTC's src/collision/environment.hh (environment.hh:12-45):
// the obstacle tables of a scene, owned by the Environment (new[] arrays,
// released by its destructor).
//
// The planner tests spheres, capsules, z-aligned capsules, cuboids and
// z-aligned cuboids.  `cylinders` is carried for source compatibility and,
// as in the reference checker (utils.cuh:211-242), is not tested: cylinders
// are meant to be stored as capsules (see factory::cylinder).
#pragma once

#include "shapes.hh"

namespace ppln::collision {

template <typename DataT>
struct Environment {
    Sphere<DataT> *spheres = nullptr;
    unsigned int num_spheres = 0;

    Capsule<DataT> *capsules = nullptr;
    unsigned int num_capsules = 0;

    Capsule<DataT> *z_aligned_capsules = nullptr;
    unsigned int num_z_aligned_capsules = 0;

    Cylinder<DataT> *cylinders = nullptr;
    unsigned int num_cylinders = 0;

    Cuboid<DataT> *cuboids = nullptr;
    unsigned int num_cuboids = 0;

    Cuboid<DataT> *z_aligned_cuboids = nullptr;
    unsigned int num_z_aligned_cuboids = 0;

    Environment() = default;
    Environment(const Environment &) = delete;
    Environment &operator=(const Environment &) = delete;
    // moves hand the tables over (a harness returns an Environment by value)
    Environment(Environment &&o) noexcept { *this = static_cast<Environment &&>(o); }
    Environment &operator=(Environment &&o) noexcept {
        if (this != &o) {
            this->~Environment();
            spheres = o.spheres; num_spheres = o.num_spheres;
            capsules = o.capsules; num_capsules = o.num_capsules;
            z_aligned_capsules = o.z_aligned_capsules; num_z_aligned_capsules = o.num_z_aligned_capsules;
            cylinders = o.cylinders; num_cylinders = o.num_cylinders;
            cuboids = o.cuboids; num_cuboids = o.num_cuboids;
            z_aligned_cuboids = o.z_aligned_cuboids; num_z_aligned_cuboids = o.num_z_aligned_cuboids;
            o.spheres = nullptr; o.capsules = nullptr; o.z_aligned_capsules = nullptr;
            o.cylinders = nullptr; o.cuboids = nullptr; o.z_aligned_cuboids = nullptr;
            o.num_spheres = o.num_capsules = o.num_z_aligned_capsules = 0;
            o.num_cylinders = o.num_cuboids = o.num_z_aligned_cuboids = 0;
        }
        return *this;
    }

    ~Environment() {
        delete[] spheres;
        delete[] capsules;
        delete[] z_aligned_capsules;
        delete[] cylinders;
        delete[] cuboids;
        delete[] z_aligned_cuboids;
    }
};

}  // namespace ppln::collision
