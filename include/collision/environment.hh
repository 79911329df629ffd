// Environment container with the member names of the reference
// (src/collision/environment.hh:11-75): six (pointer, count) pairs and an
// owns_memory flag; an owning environment delete[]s its arrays.  Counts start
// at zero here (the reference leaves them uninitialised).
#pragma once

#include "shapes.hh"

namespace ppln::collision
{
    template <typename DataT>
    struct Environment
    {
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

        bool owns_memory = true;

        Environment() = default;
        Environment(const Environment &) = delete;
        Environment &operator=(const Environment &) = delete;

        Environment(Environment &&o) noexcept
          : spheres(o.spheres), num_spheres(o.num_spheres), capsules(o.capsules), num_capsules(o.num_capsules),
            z_aligned_capsules(o.z_aligned_capsules), num_z_aligned_capsules(o.num_z_aligned_capsules),
            cylinders(o.cylinders), num_cylinders(o.num_cylinders), cuboids(o.cuboids),
            num_cuboids(o.num_cuboids), z_aligned_cuboids(o.z_aligned_cuboids),
            num_z_aligned_cuboids(o.num_z_aligned_cuboids), owns_memory(o.owns_memory)
        {
            o.spheres = nullptr;
            o.capsules = nullptr;
            o.z_aligned_capsules = nullptr;
            o.cylinders = nullptr;
            o.cuboids = nullptr;
            o.z_aligned_cuboids = nullptr;
            o.owns_memory = false;
        }

        ~Environment()
        {
            if (owns_memory) {
                delete[] spheres;
                delete[] capsules;
                delete[] z_aligned_capsules;
                delete[] cylinders;
                delete[] cuboids;
                delete[] z_aligned_cuboids;
            }
        }
    };
}  // namespace ppln::collision
