// Robot traits with the names of the reference (src/Robots.hh:9-188):
// name, dimension, Configuration, the joint-limit scale get_s_m/get_s_a and
// scale_cfg (q = u * s_m + s_a).  Host-only; `bcc_id` selects the robot model
// across the C ABI (include/bcc.h).
#pragma once

#include <array>
#include <cstddef>
#include <iostream>

#include "bcc.h"

namespace ppln::robots
{
    template <class Derived, int Dim>
    struct RobotBase
    {
        static constexpr auto dimension = Dim;
        using Configuration = std::array<float, Dim>;

        static void print_robot_config(const Configuration &cfg)
        {
            for (int i = 0; i < Dim; i++)
                std::cout << cfg[i] << ' ';
            std::cout << '\n';
        }

        static void scale_cfg(float *q)
        {
            for (int i = 0; i < Dim; ++i)
                q[i] = q[i] * Derived::get_s_m(i) + Derived::get_s_a(i);
        }
    };

    struct Panda : RobotBase<Panda, 7>
    {
        static constexpr auto name = "panda";
        static constexpr int bcc_id = BCC_PANDA;
        static constexpr float get_s_m(int i)
        {
            constexpr float v[] = {5.9342f, 3.6652f, 5.9342f, 3.2289f, 5.9342f, 3.9095999999999997f, 5.9342f};
            return v[i];
        }
        static constexpr float get_s_a(int i)
        {
            constexpr float v[] = {-2.9671f, -1.8326f, -2.9671f, -3.1416f, -2.9671f, -0.0873f, -2.9671f};
            return v[i];
        }
    };

    struct Fetch : RobotBase<Fetch, 8>
    {
        static constexpr auto name = "fetch";
        static constexpr int bcc_id = BCC_FETCH;
        static constexpr float get_s_m(int i)
        {
            constexpr float v[] = {0.38615f, 3.2112f, 2.739f, 6.28318f, 4.502f, 6.28318f, 4.32f, 6.28318f};
            return v[i];
        }
        static constexpr float get_s_a(int i)
        {
            constexpr float v[] = {0.0f, -1.6056f, -1.221f, -3.14159f, -2.251f, -3.14159f, -2.16f, -3.14159f};
            return v[i];
        }
    };

    struct Baxter : RobotBase<Baxter, 14>
    {
        static constexpr auto name = "baxter";
        static constexpr int bcc_id = BCC_BAXTER;
        static constexpr float get_s_m(int i)
        {
            constexpr float v[] = {3.40335987756f, 3.194f, 6.10835987756f, 2.6679999999999997f, 6.118f,
                                   3.66479632679f, 6.118f, 3.40335987756f, 3.194f, 6.10835987756f,
                                   2.6679999999999997f, 6.118f, 3.66479632679f, 6.118f};
            return v[i];
        }
        static constexpr float get_s_a(int i)
        {
            constexpr float v[] = {-1.70167993878f, -2.147f, -3.05417993878f, -0.05f, -3.059f, -1.57079632679f,
                                   -3.059f, -1.70167993878f, -2.147f, -3.05417993878f, -0.05f, -3.059f,
                                   -1.57079632679f, -3.059f};
            return v[i];
        }
    };

    // The reference's traits declare dimension 7 (src/Robots.hh:190-193), but
    // its generated model fkcc<Xarm7> reads q[0..12] (src/utils.cuh:38243-44112:
    // seven arm joints and the six gripper joints) and its scale tables hold
    // 13 entries (src/Robots.hh:196-203).  The model here follows the
    // generated code: 13 joints, limits from those tables.
    struct Xarm7 : RobotBase<Xarm7, 13>
    {
        static constexpr auto name = "xarm7";
        static constexpr int bcc_id = BCC_XARM7;
        static constexpr float get_s_m(int i)
        {
            constexpr float v[] = {5.6f, 3.4f, 5.6f, 3.0f, 5.6f, 3.1f, 5.6f, 0.85f, 0.85f, 0.85f, 0.85f, 0.85f, 0.85f};
            return v[i];
        }
        static constexpr float get_s_a(int i)
        {
            constexpr float v[] = {-2.8f, -1.7f, -2.8f, 0.0f, -2.8f, 0.0f, -2.8f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            return v[i];
        }
    };
}  // namespace ppln::robots
