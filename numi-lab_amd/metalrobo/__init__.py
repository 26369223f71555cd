"""MI355X-native batched temporalCone stepping (drop-in for the reference's
python/metalrobo stepping surface). Physics runs in the native library
libmetalrobo_amd.so; this package is the ctypes boundary."""

from .native import (  # noqa: F401
    BODY_STATE_DTYPE,
    CONTACT_STATUS_DTYPE,
    LEARNING_TRANSITION_DTYPE,
    MANIFOLD_HEADER_DTYPE,
    MANIFOLD_POINT_DTYPE,
    MR_RUN_SOURCE_SYNTHETIC_CHAIN_BOXES,
    MR_RUN_SOURCE_SYNTHETIC_HUMANOID_TERRAIN,
    MR_RUN_SOURCE_SYNTHETIC_RANDOM_TREES,
    MR_SYNTHETIC_CHAIN_BOXES,
    MR_SYNTHETIC_HUMANOID_TERRAIN,
    MR_SOLVER_ORDERED_SWEEPS,
    MR_SOLVER_RELAXED_JACOBI,
    MR_SYNTHETIC_RANDOM_TREES,
    PAIR_RECORD_DTYPE,
    STATUS_NAMES,
    WORLD_STATUS_DTYPE,
    TASK_ACTION_DTYPE,
    TASK_OBSERVATION_DTYPE,
    TASK_REWARD_DTYPE,
    TASK_STATE_DTYPE,
    TASK_TERMINATION_DTYPE,
    TASK_TRANSITION_DTYPE,
    LocomotionTask,
    MetalRoboError,
    TaskRollout,
    World,
    WorldModel,
    WorldStepConfigC,
    load,
    version,
)
from .packs import (  # noqa: F401
    DenseLayer,
    Policy,
    PolicyPack,
    learning_pack_content_hash,
    read_policy_rollout_pack,
    write_policy_rollout_pack,
)
