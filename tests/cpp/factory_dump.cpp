// Prints the drop-in factory's obstacles for fixed inputs, one per line.
#include <cstdio>

#include "src/collision/factory.hh"

using namespace ppln::collision;

int main() {
    const float euler[3][3] = {{0.3f, -0.7f, 1.9f}, {0.f, 0.f, 0.f}, {1.2f, 0.4f, -2.5f}};
    for (auto &e : euler) {
        auto c = factory::cuboid::array({0.1f, 0.2f, 0.3f}, {e[0], e[1], e[2]}, {0.4f, 0.5f, 0.6f});
        std::printf("cuboid %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", c.x, c.y, c.z,
                    c.axis_1_x, c.axis_1_y, c.axis_1_z, c.axis_2_x, c.axis_2_y, c.axis_2_z, c.axis_3_x, c.axis_3_y,
                    c.axis_3_z, c.axis_1_r, c.axis_2_r, c.axis_3_r);
        auto k = factory::cylinder::center::array({0.1f, 0.2f, 0.3f}, {e[0], e[1], e[2]}, 0.05f, 0.4f);
        std::printf("capsule %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", k.x1, k.y1, k.z1, k.xv, k.yv, k.zv, k.r, k.rdv);
    }
    return 0;
}
