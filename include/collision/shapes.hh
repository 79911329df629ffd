// Obstacle primitives with the field names and order of the reference's data
// model (src/collision/shapes.hh:12-245), so code written against the
// reference compiles unchanged.  Only the numeric fields reach the checker;
// `name` and `min_distance` are carried for source compatibility.
#pragma once

#include <cmath>
#include <string>

namespace ppln::collision
{
    template <typename DataT>
    struct Shape
    {
        std::string name;
        DataT min_distance{};

        Shape() = default;

        template <typename OtherDataT>
        explicit Shape(const Shape<OtherDataT> &other) : name(other.name), min_distance(other.min_distance)
        {
        }
    };

    // Cuboid: centre, three axes (columns of its rotation), half-extents
    // along each axis (src/collision/shapes.hh:26-126).
    template <typename DataT>
    struct Cuboid : public Shape<DataT>
    {
        DataT x, y, z;
        DataT axis_1_x, axis_1_y, axis_1_z;
        DataT axis_2_x, axis_2_y, axis_2_z;
        DataT axis_3_x, axis_3_y, axis_3_z;
        DataT axis_1_r, axis_2_r, axis_3_r;

        Cuboid() = default;

        explicit Cuboid(DataT x, DataT y, DataT z, DataT a1x, DataT a1y, DataT a1z, DataT a2x, DataT a2y, DataT a2z,
                        DataT a3x, DataT a3y, DataT a3z, DataT r1, DataT r2, DataT r3)
          : x(x), y(y), z(z), axis_1_x(a1x), axis_1_y(a1y), axis_1_z(a1z), axis_2_x(a2x), axis_2_y(a2y),
            axis_2_z(a2z), axis_3_x(a3x), axis_3_y(a3y), axis_3_z(a3z), axis_1_r(r1), axis_2_r(r2), axis_3_r(r3)
        {
            // distance from the origin to the box surface/interior
            const DataT d[3] = {-(x * a1x + y * a1y + z * a1z), -(x * a2x + y * a2y + z * a2z),
                                -(x * a3x + y * a3y + z * a3z)};
            const DataT h[3] = {r1, r2, r3};
            DataT s = 0;
            for (int k = 0; k < 3; ++k) {
                const DataT e = std::fabs(d[k]) - h[k];
                if (e > 0)
                    s += e * e;
            }
            this->min_distance = std::sqrt(s);
        }
    };

    // Cylinder / capsule: first end point, vector to the second end point,
    // radius, and rdv = 1 / (v . v) (src/collision/shapes.hh:128-222).
    template <typename DataT>
    struct Cylinder : public Shape<DataT>
    {
        DataT x1, y1, z1;
        DataT xv, yv, zv;
        DataT r;
        DataT rdv;

        Cylinder() = default;

        explicit Cylinder(DataT x1, DataT y1, DataT z1, DataT xv, DataT yv, DataT zv, DataT r, DataT rdv)
          : x1(x1), y1(y1), z1(z1), xv(xv), yv(yv), zv(zv), r(r), rdv(rdv)
        {
            DataT t = -(x1 * xv + y1 * yv + z1 * zv) * rdv;
            t = t < 0 ? 0 : (t > 1 ? 1 : t);
            const DataT px = x1 + xv * t, py = y1 + yv * t, pz = z1 + zv * t;
            const DataT d = std::sqrt(px * px + py * py + pz * pz) - r;
            this->min_distance = d > 0 ? d : 0;
        }

        DataT x2() const { return x1 + xv; }
        DataT y2() const { return y1 + yv; }
        DataT z2() const { return z1 + zv; }
    };

    template <typename DataT>
    using Capsule = Cylinder<DataT>;

    template <typename DataT>
    struct Sphere : public Shape<DataT>
    {
        DataT x, y, z, r;

        Sphere() = default;

        explicit Sphere(DataT x, DataT y, DataT z, DataT r) : x(x), y(y), z(z), r(r)
        {
            this->min_distance = std::sqrt(x * x + y * y + z * z) - r;
        }
    };
}  // namespace ppln::collision
