// Internals shared by the host translation units of libbcc (not part of the
// C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bcc.h"
#include "bcc_dense.h"
#include "robots/bcc_model.h"

struct bcc_scene {
    int num_envs = 0;
    float* d_obs = nullptr;      // all environments, device layout (see bcc_device.h)
    int* d_meta = nullptr;       // [num_envs][kEnvMetaInts]
    float* d_ref_obs = nullptr;  // the reference's lists and order, for the work counter
    int* d_ref_meta = nullptr;
    size_t obs_floats = 0;
    std::vector<int> obstacles;  // tested primitives per environment (host copy)
    int cull_min = 0;            // culling threshold stamped into every environment's record
};

namespace bcc_detail {

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
const bcc_robot_model* host_model(int robot);
int device_model(int robot, bcc::ModelDev* out);

// Robot a model was created for.
int model_robot(const bcc_model* model);

// Inference operands of `model` for the environments of `scene`, in the
// model's own buffers: packs the weight fragments and encodes the scene on
// stream s (allocating the code buffer, after a synchronisation, only when
// the scene has more environments than any before).
int model_mlp_args(bcc_model* model, const bcc_scene* scene, hipStream_t s, bcc::MlpArgs* out);

// The same operands as one device allocation *owned (hipFree it), filled and
// synchronised: a snapshot of the model as it is now, for plans.
int model_mlp_snapshot(const bcc_model* model, const bcc_scene* scene, bcc::MlpArgs* out, void** owned);

// Lower limit and range of joint j of a robot (the model's feature scaling).
void joint_range(int robot, int j, float* lo, float* range);

}  // namespace bcc_detail

#define BCC_HIP(call, what)                                  \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess)                                \
            return bcc_detail::hip_fail(_e, what);           \
    } while (0)
