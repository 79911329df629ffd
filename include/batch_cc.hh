// Drop-in for the reference's entry point (src/batch_cc.hh:4-7):
//
//   template <typename Robot>
//   void batch_cc::batch_cc(std::vector<Environment<float>>& h_envs,
//                           std::vector<std::array<Robot::Configuration, 2>>& edges,
//                           int resolution, std::vector<bool>& results);
//
// Same arguments and result layout (results[edge * E + env], true = the edge
// collides in that environment, src/batch_cc.cu:131,319-321); the work runs
// on the gfx950 kernels behind include/bcc.h.  Errors throw
// std::runtime_error, the reference's exception type for setup errors
// (src/batch_cc.cu:25-32,45-56).  `results` is resized to E*M if it is
// smaller (the reference requires the caller to pre-size it).  Nothing is
// printed unless BCC_VERBOSE is set in the environment.
#pragma once

#include <array>
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "Robots.hh"
#include "bcc.h"
#include "collision/environment.hh"

namespace batch_cc
{
    namespace detail
    {
        inline void check(int rc)
        {
            if (rc != BCC_OK)
                throw std::runtime_error(std::string("batch_cc: ") + bcc_last_error());
        }

        // Numeric fields of each primitive, in the C ABI's flat layout.
        inline void pack(const ppln::collision::Environment<float> &e, std::vector<float> *buf, bcc_env_desc &d)
        {
            const bool ok = (e.num_spheres == 0 || e.spheres) && (e.num_capsules == 0 || e.capsules) &&
                            (e.num_z_aligned_capsules == 0 || e.z_aligned_capsules) &&
                            (e.num_cylinders == 0 || e.cylinders) && (e.num_cuboids == 0 || e.cuboids) &&
                            (e.num_z_aligned_cuboids == 0 || e.z_aligned_cuboids);
            if (!ok)
                throw std::runtime_error("Host environment has positive counts with null pointers.");
            for (unsigned i = 0; i < e.num_spheres; ++i) {
                const auto &s = e.spheres[i];
                buf[0].insert(buf[0].end(), {s.x, s.y, s.z, s.r});
            }
            auto cyl = [](std::vector<float> &b, const ppln::collision::Cylinder<float> &c) {
                b.insert(b.end(), {c.x1, c.y1, c.z1, c.xv, c.yv, c.zv, c.r, c.rdv});
            };
            auto cub = [](std::vector<float> &b, const ppln::collision::Cuboid<float> &c) {
                b.insert(b.end(), {c.x, c.y, c.z, c.axis_1_x, c.axis_1_y, c.axis_1_z, c.axis_2_x, c.axis_2_y,
                                   c.axis_2_z, c.axis_3_x, c.axis_3_y, c.axis_3_z, c.axis_1_r, c.axis_2_r,
                                   c.axis_3_r});
            };
            for (unsigned i = 0; i < e.num_capsules; ++i)
                cyl(buf[1], e.capsules[i]);
            for (unsigned i = 0; i < e.num_z_aligned_capsules; ++i)
                cyl(buf[2], e.z_aligned_capsules[i]);
            for (unsigned i = 0; i < e.num_cylinders; ++i)
                cyl(buf[3], e.cylinders[i]);
            for (unsigned i = 0; i < e.num_cuboids; ++i)
                cub(buf[4], e.cuboids[i]);
            for (unsigned i = 0; i < e.num_z_aligned_cuboids; ++i)
                cub(buf[5], e.z_aligned_cuboids[i]);
            d = bcc_env_desc{buf[0].data(), e.num_spheres,           buf[1].data(), e.num_capsules,
                             buf[2].data(), e.num_z_aligned_capsules, buf[3].data(), e.num_cylinders,
                             buf[4].data(), e.num_cuboids,           buf[5].data(), e.num_z_aligned_cuboids};
        }

        struct SceneHandle
        {
            bcc_scene *s = nullptr;
            ~SceneHandle() { bcc_scene_destroy(s); }
        };
    }  // namespace detail

    template <typename Robot>
    void batch_cc(std::vector<ppln::collision::Environment<float>> &h_envs,
                  std::vector<std::array<typename Robot::Configuration, 2>> &edges, int resolution,
                  std::vector<bool> &results)
    {
        const bool verbose = std::getenv("BCC_VERBOSE") != nullptr;
        const auto t0 = std::chrono::steady_clock::now();
        const size_t E = h_envs.size();
        const size_t M = edges.size();
        std::vector<std::vector<float>> store(E * 6);
        std::vector<bcc_env_desc> descs(E);
        for (size_t e = 0; e < E; ++e)
            detail::pack(h_envs[e], &store[e * 6], descs[e]);
        detail::SceneHandle scene;
        detail::check(bcc_scene_create(descs.data(), (int)E, &scene.s));
        if (results.size() < E * M)
            results.resize(E * M);
        if (E == 0 || M == 0)
            return;
        std::vector<uint8_t> out(E * M);
        static_assert(sizeof(std::array<typename Robot::Configuration, 2>) == sizeof(float) * 2 * Robot::dimension,
                      "edges must be contiguous [M][2][dim] floats");
        const auto t1 = std::chrono::steady_clock::now();
        detail::check(bcc_check_dense(Robot::bcc_id, scene.s, reinterpret_cast<const float *>(edges.data()),
                                      (int)M, resolution, out.data()));
        const auto t2 = std::chrono::steady_clock::now();
        for (size_t i = 0; i < E * M; ++i)
            results[i] = out[i] != 0;
        if (verbose) {
            const double setup = std::chrono::duration<double>(t1 - t0).count();
            const double check = std::chrono::duration<double>(t2 - t1).count();
            std::cout << "Setup time: " << setup << " s\n"
                      << "Check time (incl. copies): " << check << " s\n"
                      << "Edges checked: " << E * M << "\n"
                      << "Throughput: " << (double)(E * M) / check << " edges/s\n";
        }
    }
}  // namespace batch_cc
