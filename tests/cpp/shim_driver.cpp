// Calls the drop-in entry point exactly as the reference's README driver does
// (scripts/batch_cc_mbm.cpp:125-162): build environments with the reference
// data model, sample edges, call batch_cc::batch_cc<Robot>, print the
// edge-major verdict matrix.
//
//   shim_driver <robot> <scene.bin> <edges.bin>   -> prints "E M" then M lines of E 0/1
//   shim_driver nullcheck                          -> exercises the error path
//
// scene.bin: int32 E, then per env int32 counts[6] followed by the floats of
// each kind (4, 8, 8, 8, 15, 15 per primitive).  edges.bin: int32 M, int32
// dim, then M*2*dim floats.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "batch_cc.hh"

using namespace ppln::collision;

static std::vector<Environment<float>> read_scene(const char *path)
{
    std::ifstream f(path, std::ios::binary);
    int E = 0;
    f.read((char *)&E, 4);
    std::vector<Environment<float>> envs;
    envs.reserve(E);
    for (int e = 0; e < E; ++e) {
        int c[6];
        f.read((char *)c, sizeof c);
        Environment<float> env;
        auto rd = [&](int n, int stride) {
            std::vector<float> v((size_t)n * stride);
            f.read((char *)v.data(), sizeof(float) * v.size());
            return v;
        };
        auto s = rd(c[0], 4);
        if (c[0]) {
            env.spheres = new Sphere<float>[c[0]];
            for (int i = 0; i < c[0]; ++i)
                env.spheres[i] = Sphere<float>(s[4 * i], s[4 * i + 1], s[4 * i + 2], s[4 * i + 3]);
            env.num_spheres = c[0];
        }
        auto cyl = [&](int n, Cylinder<float> *&dst, unsigned &cnt) {
            auto v = rd(n, 8);
            if (!n)
                return;
            dst = new Cylinder<float>[n];
            for (int i = 0; i < n; ++i) {
                const float *p = &v[8 * i];
                dst[i] = Cylinder<float>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7]);
            }
            cnt = n;
        };
        cyl(c[1], env.capsules, env.num_capsules);
        cyl(c[2], env.z_aligned_capsules, env.num_z_aligned_capsules);
        cyl(c[3], env.cylinders, env.num_cylinders);
        auto cub = [&](int n, Cuboid<float> *&dst, unsigned &cnt) {
            auto v = rd(n, 15);
            if (!n)
                return;
            dst = new Cuboid<float>[n];
            for (int i = 0; i < n; ++i) {
                const float *p = &v[15 * i];
                dst[i] = Cuboid<float>(p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                                       p[12], p[13], p[14]);
            }
            cnt = n;
        };
        cub(c[4], env.cuboids, env.num_cuboids);
        cub(c[5], env.z_aligned_cuboids, env.num_z_aligned_cuboids);
        envs.push_back(std::move(env));
    }
    return envs;
}

template <class Robot>
static int run(const char *scene, const char *edge_path)
{
    auto envs = read_scene(scene);
    std::ifstream f(edge_path, std::ios::binary);
    int M = 0, dim = 0;
    f.read((char *)&M, 4);
    f.read((char *)&dim, 4);
    if (dim != Robot::dimension)
        throw std::runtime_error("edge dimension mismatch");
    std::vector<std::array<typename Robot::Configuration, 2>> edges(M);
    f.read((char *)edges.data(), sizeof(float) * 2 * dim * (size_t)M);
    std::vector<bool> results(envs.size() * (size_t)M);
    batch_cc::batch_cc<Robot>(envs, edges, 32, results);
    std::cout << envs.size() << " " << M << "\n";
    for (int i = 0; i < M; ++i) {
        for (size_t j = 0; j < envs.size(); ++j)
            std::cout << results[i * envs.size() + j] << (j + 1 < envs.size() ? " " : "");
        std::cout << "\n";
    }
    return 0;
}

int main(int argc, char **argv)
{
    try {
        if (argc == 2 && std::string(argv[1]) == "nullcheck") {
            std::vector<Environment<float>> envs(1);
            envs[0].num_spheres = 3;   // positive count, null pointer
            std::vector<std::array<ppln::robots::Panda::Configuration, 2>> edges(1);
            std::vector<bool> results(1);
            try {
                batch_cc::batch_cc<ppln::robots::Panda>(envs, edges, 32, results);
            } catch (const std::runtime_error &e) {
                std::cout << "runtime_error: " << e.what() << "\n";
                return 0;
            }
            std::cout << "no exception\n";
            return 1;
        }
        if (argc != 4)
            return 2;
        const std::string robot = argv[1];
        if (robot == "panda")
            return run<ppln::robots::Panda>(argv[2], argv[3]);
        if (robot == "fetch")
            return run<ppln::robots::Fetch>(argv[2], argv[3]);
        if (robot == "baxter")
            return run<ppln::robots::Baxter>(argv[2], argv[3]);
        return 2;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << "\n";
        return 3;
    }
}
