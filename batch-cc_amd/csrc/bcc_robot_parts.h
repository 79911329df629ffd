// Kernels of one robot, split into translation units that compile in
// parallel: the generated check program is large (straight-line code), and
// each kernel variant that instantiates it costs minutes of compile time.
//
// Included at the end of robot_<name>.hip after
//   BCC_ROBOT       robot name (panda, fetch, baxter)
//   BCC_ROBOT_T     its check struct (Robot_<name>)
//   BCC_KNOBS       HW, HU, LW, LU: waves per SIMD and obstacle-loop unroll of
//                   the high- and low-occupancy kernel variants
// BCC_PART selects what this translation unit instantiates:
//   unset  everything
//   0      host dispatchers and the small kernels (FK, samples, work, labels,
//          clearance audit)
//   1, 7   ordered dense kernel, low / high occupancy
//   2, 3   plain dense kernel, high / low occupancy
//   4, 5   ragged kernel, high / low occupancy
//   6, 8   learned-order ragged kernel (predicted probe + plain finish), low / high occupancy
//   9, 10  fused learned kernel (probe, inference, probe, plain finish), low / high occupancy
#ifndef BCC_PART
#define BCC_PART -1
#endif

namespace bcc {
namespace robot_parts {

using RobotT = BCC_ROBOT_T;
constexpr VariantKnobs kKnobs = {BCC_KNOBS};
constexpr int HW = kKnobs.hi_waves, HU = kKnobs.hi_unroll, LW = kKnobs.lo_waves, LU = kKnobs.lo_unroll;

}  // namespace robot_parts

#define BCC_CAT2_(a, b) a##b
#define BCC_CAT2(a, b) BCC_CAT2_(a, b)

#if BCC_PART == 0
// instantiated in parts 1-8
extern template hipError_t launch_dense_variant<robot_parts::RobotT, true, robot_parts::LW, robot_parts::LU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
extern template hipError_t launch_dense_variant<robot_parts::RobotT, true, robot_parts::HW, robot_parts::HU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
extern template hipError_t launch_dense_variant<robot_parts::RobotT, false, robot_parts::HW, robot_parts::HU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
extern template hipError_t launch_dense_variant<robot_parts::RobotT, false, robot_parts::LW, robot_parts::LU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
extern template hipError_t launch_ragged_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(
    const RaggedArgs&, hipStream_t);
extern template hipError_t launch_ragged_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(
    const RaggedArgs&, hipStream_t);
extern template hipError_t launch_learn_finish_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(
    const LearnArgs&, hipStream_t);
extern template hipError_t launch_learn_finish_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(
    const LearnArgs&, hipStream_t);
extern template hipError_t launch_learned_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(
    const LearnedArgs&, hipStream_t);
extern template hipError_t launch_learned_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(
    const LearnedArgs&, hipStream_t);
#endif

#if BCC_PART == -1 || BCC_PART == 0
const VariantKnobs BCC_CAT2(knobs_, BCC_ROBOT) = robot_parts::kKnobs;

hipError_t BCC_CAT2(launch_dense_, BCC_ROBOT)(const DenseArgs& a, hipStream_t stream)
{
    using namespace robot_parts;
    return launch_dense<RobotT, HW, HU, LW, LU>(a, stream);
}

hipError_t BCC_CAT2(launch_ragged_, BCC_ROBOT)(const RaggedArgs& a, hipStream_t stream)
{
    using namespace robot_parts;
    return launch_ragged<RobotT, HW, HU, LW, LU>(a, stream);
}

hipError_t BCC_CAT2(launch_learn_, BCC_ROBOT)(const LearnArgs& a, hipStream_t stream)
{
    using namespace robot_parts;
    return launch_learn<RobotT, HW, HU, LW, LU>(a, stream);
}

hipError_t BCC_CAT2(launch_learned_, BCC_ROBOT)(const LearnedArgs& a, hipStream_t stream)
{
    using namespace robot_parts;
    return launch_learned<RobotT, HW, HU, LW, LU>(a, stream);
}

hipError_t BCC_CAT2(launch_fk_, BCC_ROBOT)(const float* qs, int n, float* out, hipStream_t stream)
{
    return launch_fk<robot_parts::RobotT>(qs, n, out, stream);
}

hipError_t BCC_CAT2(launch_samples_, BCC_ROBOT)(const float* edges, int m, int resolution, int* n_out, float* cfg0,
                                                float* cfg_last, hipStream_t stream)
{
    return launch_samples<robot_parts::RobotT>(edges, m, resolution, n_out, cfg0, cfg_last, stream);
}

hipError_t BCC_CAT2(launch_work_, BCC_ROBOT)(const WorkArgs& a, hipStream_t stream)
{
    return launch_work<robot_parts::RobotT>(a, stream);
}

hipError_t BCC_CAT2(launch_label_, BCC_ROBOT)(const LabelArgs& a, hipStream_t stream)
{
    return launch_label<robot_parts::RobotT>(a, stream);
}

hipError_t BCC_CAT2(launch_audit_, BCC_ROBOT)(const AuditArgs& a, hipStream_t stream)
{
    return launch_audit<robot_parts::RobotT>(a, stream);
}
#endif

#if BCC_PART == 1
template hipError_t launch_dense_variant<robot_parts::RobotT, true, robot_parts::LW, robot_parts::LU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
#endif
#if BCC_PART == 7
template hipError_t launch_dense_variant<robot_parts::RobotT, true, robot_parts::HW, robot_parts::HU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
#endif
#if BCC_PART == 2
template hipError_t launch_dense_variant<robot_parts::RobotT, false, robot_parts::HW, robot_parts::HU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
#endif
#if BCC_PART == 3
template hipError_t launch_dense_variant<robot_parts::RobotT, false, robot_parts::LW, robot_parts::LU>(
    const DenseArgs&, unsigned, int, int, hipStream_t);
#endif
#if BCC_PART == 4
template hipError_t launch_ragged_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(const RaggedArgs&,
                                                                                                 hipStream_t);
#endif
#if BCC_PART == 5
template hipError_t launch_ragged_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(const RaggedArgs&,
                                                                                                 hipStream_t);
#endif
#if BCC_PART == 6
template hipError_t launch_learn_finish_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(
    const LearnArgs&, hipStream_t);
#endif
#if BCC_PART == 8
template hipError_t launch_learn_finish_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(
    const LearnArgs&, hipStream_t);
#endif
#if BCC_PART == 9
template hipError_t launch_learned_variant<robot_parts::RobotT, robot_parts::LW, robot_parts::LU>(const LearnedArgs&,
                                                                                                  hipStream_t);
#endif
#if BCC_PART == 10
template hipError_t launch_learned_variant<robot_parts::RobotT, robot_parts::HW, robot_parts::HU>(const LearnedArgs&,
                                                                                                  hipStream_t);
#endif

}  // namespace bcc
