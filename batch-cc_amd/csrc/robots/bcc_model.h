/* Robot model as plain data: kinematic tree, sphere table and the ordered
 * collision-check program.  One instance per robot is generated into
 * <robot>_model.h by batch-cc_amd/robots/gen_robot_code.py.
 *
 * The check program mirrors the nesting of the reference's generated
 * fkcc<Robot> (src/utils.cuh:342-5442 Panda, :10382-24992 Fetch,
 * :24994-38241 Baxter):
 *   env group g : gate sphere env_gate[g]; if env_lone[g] the gate alone
 *                 decides, otherwise the fine spheres
 *                 env_fine[env_fine_off[g] .. env_fine_off[g+1]) are tested
 *                 only when the gate hits (src/utils.cuh:814-1173);
 *   self group g: gate pair (self_gate_a[g], self_gate_b[g]); fine pairs are
 *                 the cross product self_a[..] x self_b[..]
 *                 (src/utils.cuh:1197-5434).
 * prog_kind/prog_index give the reference's program order (0 = env, 1 =
 * self); prog_fk_ops/prog_fk_trig count the reference's source-level FK
 * arithmetic executed before each item (used only for work accounting).
 */
#ifndef BCC_MODEL_H
#define BCC_MODEL_H

typedef struct bcc_robot_model {
    const char *name;
    int dim;
    int num_spheres;
    int num_env_groups;
    int num_self_groups;
    int num_program;
    const int *joint_order;      /* evaluation order of joints (parents first) */
    const int *joint_parent;     /* -1 = robot base frame */
    const int *joint_type;       /* 0 revolute about local z, 1 prismatic along local z */
    const float (*joint_R)[9];   /* fixed rotation parent->joint at q=0, row-major */
    const float (*joint_t)[3];   /* fixed translation in the parent frame */
    const double (*joint_R64)[9];
    const double (*joint_t64)[3];
    const int *sphere_frame;     /* joint frame the sphere is attached to, -1 = base */
    const float (*sphere_p)[3];  /* local centre */
    const double (*sphere_p64)[3];
    const float *sphere_r;       /* float32 radius (unpadded) */
    const int *env_gate;
    const int *env_lone;
    const int *env_fine_off;
    const int *env_fine;
    const int *self_gate_a;
    const int *self_gate_b;
    const int *self_lone;
    const int *self_a_off;
    const int *self_a;
    const int *self_b_off;
    const int *self_b;
    const int *prog_kind;
    const int *prog_index;
    const int *prog_fk_ops;
    const int *prog_fk_trig;
    int ref_fk_flops;            /* reference source-level FK binary ops per sample */
    int ref_fk_trig;             /* reference sin/cos calls per sample */
} bcc_robot_model;

#endif /* BCC_MODEL_H */
